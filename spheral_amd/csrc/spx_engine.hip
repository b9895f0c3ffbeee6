// spx_engine.hip -- host orchestration and C ABI of the MI355X SPH engine.
//
// One spx_run = one FluidNodeList evolved by SPH (+ MonaghanGingold Q +
// SPHSmoothingScale, + TreeGravity when enabled) under CheapSynchronousRK2, with
// all state resident in HBM.  The host enqueues kernels on the run's stream in
// the reference's call order (Integrator/Integrator.cc:67-111,
// Integrator/CheapSynchronousRK2.cc:40-132) and reads back only the grid extent
// before the neighbour build and dt/overflow words after it.
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/spx_abi.h"
#include "spx_error.hpp"
#include "spx_gravity.hpp"
#include "spx_sph.hpp"
#include "spx_table.hpp"

namespace spx {

static thread_local std::string g_last_error;

static inline int nblocks(int64_t n) { return (int)((n + kBlock - 1)/kBlock); }

__global__ void k_advance_time(Scalars* sc) {
  sc->t = sc->t + sc->dt;
  sc->lastDt = sc->dt;
  sc->cycle = sc->cycle + 1;
}

class Run {
 public:
  spx_config cfg;
  int D;
  int64_t n;
  int device;
  hipStream_t stream = nullptr;
  KernelTable table;
  DevTable dtab{};
  Phys ph{};
  Geo geo{};
  Fields f{};
  Fields alt{};  // reorder targets for the carried arrays
  GravityTree grav;

  std::vector<void*> allocs;
  double* slab = nullptr;
  double* slabAlt = nullptr;
  uint64_t *keys = nullptr, *keysAlt = nullptr;
  uint32_t *vals = nullptr, *valsAlt = nullptr;
  void* sortTmp = nullptr;
  size_t sortTmpBytes = 0;
  int* cellStart = nullptr;
  int* cellEnd = nullptr;
  int64_t cellsAlloc = 0;
  double* part = nullptr;        // per-block partials
  Scalars* dsc = nullptr;        // device scalars
  Scalars* hsc = nullptr;        // pinned mirror
  double* hpart = nullptr;       // pinned partials
  int64_t* oidAlt = nullptr;
  int ld = 0;
  bool haveState = false, initialized = false, connValid = false, wrapPending = false;
  int64_t stepsSinceReorder = 0;
  double t = 0.0, lastDt = 0.0;
  int64_t cycle = 0;
  int64_t npairsCache = -1;
  // kernel timing (ms) of the last step per class
  std::vector<std::string> tnames;
  std::vector<double> tms;
  std::vector<hipEvent_t> evs;
  bool timing = false;

  Run(const spx_config& c, int64_t n_, int dev) : cfg(c), D(c.dim), n(n_), device(dev) {
    validate();
    HIP_CHECK(hipSetDevice(device));
    HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    table.build(cfg.kernel, D, (unsigned)cfg.kernel_points, cfg.min_nperh, cfg.max_nperh);
    uploadTable();
    setPhys();
    ld = (int)((n + 255)/256*256);
    allocate();
    t = 0.0;
    lastDt = cfg.last_dt;
    grav.init(D, n, ld, stream, cfg.gravity != 0);
  }
  ~Run() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto e : evs) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
    if (hsc) (void)hipHostFree(hsc);
    if (hpart) (void)hipHostFree(hpart);
    grav.release();
    if (stream) (void)hipStreamDestroy(stream);
  }

  void validate() {
    if (cfg.dim < 1 || cfg.dim > 3) raise(SPX_ERR_ARG, "dim must be 1, 2 or 3");
    if (n < 1 || n > (int64_t)kIdxMask) raise(SPX_ERR_CAPACITY, "node count must be in [1, 2^26)");
    if (cfg.kernel != SPX_KERNEL_BSPLINE && cfg.kernel != SPX_KERNEL_WENDLANDC4) raise(SPX_ERR_ARG, "unknown kernel");
    if (cfg.kernel_points < 3) raise(SPX_ERR_ARG, "kernel_points must be >= 3");
    if (!cfg.compatible_energy) raise(SPX_ERR_UNSUPPORTED, "only compatibleEnergyEvolution=True is implemented");
    if (cfg.precision != SPX_PREC_FP64 && cfg.precision != SPX_PREC_MIXED) raise(SPX_ERR_ARG, "unknown precision");
    if (cfg.neighbor_extent <= 0.0) raise(SPX_ERR_ARG, "neighbor_extent must be positive");
    for (int a = 0; a < 3; ++a) {
      const bool per = cfg.bc_lo[a] == SPX_BC_PERIODIC || cfg.bc_hi[a] == SPX_BC_PERIODIC;
      if (per && cfg.bc_lo[a] != cfg.bc_hi[a]) raise(SPX_ERR_ARG, "periodic must be set on both sides of an axis");
      if (a < cfg.dim && (cfg.bc_lo[a] != SPX_BC_NONE || cfg.bc_hi[a] != SPX_BC_NONE) && !(cfg.xmax[a] > cfg.xmin[a]))
        raise(SPX_ERR_ARG, "boundary planes need xmax > xmin");
    }
    if (cfg.gravity && (cfg.G <= 0.0 || cfg.softening <= 0.0 || cfg.opening <= 0.0 || cfg.ftimestep <= 0.0))
      raise(SPX_ERR_ARG, "TreeGravity needs G, softening, opening, ftimestep > 0");
  }

  template <typename T>
  T* dalloc(size_t count) {
    void* p = nullptr;
    HIP_CHECK(hipMalloc(&p, std::max<size_t>(count, 1)*sizeof(T)));
    allocs.push_back(p);
    return static_cast<T*>(p);
  }

  void uploadTable() {
    const size_t nint = table.W.N1 + 1;
    std::vector<double> coef(6*nint);
    for (size_t k = 0; k < nint; ++k)
      for (int c = 0; c < 3; ++c) { coef[6*k + c] = table.W.coef[3*k + c]; coef[6*k + 3 + c] = table.gW.coef[3*k + c]; }
    double* dc = dalloc<double>(coef.size());
    double* dh = dalloc<double>(table.nperh.vals.size());
    HIP_CHECK(hipMemcpy(dc, coef.data(), coef.size()*sizeof(double), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(dh, table.nperh.vals.data(), table.nperh.vals.size()*sizeof(double), hipMemcpyHostToDevice));
    dtab.coef = dc;
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
    ph.ext2 = cfg.neighbor_extent*cfg.neighbor_extent;
    ph.hmin = cfg.hmin;
    ph.hmax = cfg.hmax;
    ph.rho_min = cfg.rho_min;
    ph.rho_max = cfg.rho_max;
    ph.hmaxInv = 1.0/cfg.hmax;
    ph.hminInv = 1.0/cfg.hmin;
    ph.correctVG = cfg.correct_velocity_gradient;
    ph.useVmagDt = cfg.use_velocity_magnitude_for_dt;
    ph.gravity = cfg.gravity;
    for (int a = 0; a < 3; ++a) {
      geo.bc_lo[a] = a < D ? cfg.bc_lo[a] : BC_NONE;
      geo.bc_hi[a] = a < D ? cfg.bc_hi[a] : BC_NONE;
      geo.xmin[a] = cfg.xmin[a];
      geo.xmax[a] = cfg.xmax[a];
      geo.nc[a] = 1;
      geo.cell[a] = 1.0;
      geo.inv_cell[a] = 1.0;
      geo.origin[a] = 0.0;
    }
    geo.permask = 0;
    geo.reflmask = 0;
    for (int a = 0; a < D; ++a) {
      if (geo.bc_lo[a] == BC_PERIODIC) geo.permask |= 1 << a;
      if (geo.bc_lo[a] == BC_REFLECTING) geo.reflmask |= 1 << (2*a);
      if (geo.bc_hi[a] == BC_REFLECTING) geo.reflmask |= 1 << (2*a + 1);
    }
  }

  // Carried arrays (persist across steps, follow the reorder).
  std::vector<double**> carried(Fields& F) {
    std::vector<double**> v;
    for (int a = 0; a < D; ++a) { v.push_back(&F.x[a]); v.push_back(&F.v[a]); }
    for (double** p : {&F.m, &F.rho, &F.eps, &F.hinv, &F.P, &F.cs, &F.omega, &F.DepsDt, &F.DrhoDt, &F.DHDt, &F.maxQ, &F.pot})
      v.push_back(p);
    for (int a = 0; a < D; ++a) { v.push_back(&F.DxDt[a]); v.push_back(&F.DvDt[a]); v.push_back(&F.DvDtG[a]); }
    return v;
  }

  void allocate() {
    // Slab of per-node arrays, each ld doubles.
    std::vector<double**> all;
    Fields& F = f;
    for (int a = 0; a < D; ++a) {
      for (double** p : {&F.x[a], &F.v[a], &F.xm[a], &F.vm[a], &F.v1[a], &F.DxDt[a], &F.DvDt[a], &F.DvDtG[a],
                         &F.gradRho[a], &F.m1[a]})
        all.push_back(p);
    }
    for (double** p : {&F.m, &F.rho, &F.eps, &F.hinv, &F.P, &F.cs, &F.omega, &F.rhom, &F.epsm, &F.hinvm, &F.Pm,
                       &F.csm, &F.omegam, &F.prhom, &F.DepsDt, &F.DrhoDt, &F.DHDt, &F.maxQ, &F.Hideal, &F.m0, &F.rhoSum,
                       &F.norm, &F.effQ, &F.ns, &F.pot})
      all.push_back(p);
    const size_t tensorComps = 2*(size_t)D*D;   // DvDx, Mcorr
    const size_t total = (all.size() + tensorComps)*(size_t)ld;
    slab = dalloc<double>(total);
    HIP_CHECK(hipMemset(slab, 0, total*sizeof(double)));
    size_t off = 0;
    for (double** p : all) { *p = slab + off; off += ld; }
    F.DvDx = slab + off; off += (size_t)D*D*ld;
    F.Mcorr = slab + off; off += (size_t)D*D*ld;
    for (int a = D; a < 3; ++a) {
      F.x[a] = F.v[a] = F.xm[a] = F.vm[a] = F.v1[a] = F.DxDt[a] = F.DvDt[a] = F.DvDtG[a] = F.gradRho[a] = F.m1[a] = nullptr;
    }
    // alternate storage for the carried arrays
    alt = f;
    auto cl = carried(alt);
    const size_t altTotal = (cl.size() + (size_t)D*D)*(size_t)ld;
    slabAlt = dalloc<double>(altTotal);
    HIP_CHECK(hipMemset(slabAlt, 0, altTotal*sizeof(double)));
    off = 0;
    for (double** p : cl) { *p = slabAlt + off; off += ld; }
    alt.DvDx = slabAlt + off;
    f.oid = dalloc<int64_t>(ld);
    oidAlt = dalloc<int64_t>(ld);
    f.cnt = dalloc<int32_t>(ld);
    HIP_CHECK(hipMemset(f.cnt, 0, ld*sizeof(int32_t)));
    f.ld = ld;
    f.cap = 0;
    f.nbr = nullptr;
    keys = dalloc<uint64_t>(ld);
    keysAlt = dalloc<uint64_t>(ld);
    vals = dalloc<uint32_t>(ld);
    valsAlt = dalloc<uint32_t>(ld);
    size_t tmp = 0;
    HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp, keys, keysAlt, vals, valsAlt, (size_t)n, 0, 63, stream));
    sortTmpBytes = tmp;
    sortTmp = dalloc<char>(tmp);
    part = dalloc<double>((size_t)nblocks(n)*(1 + 2*3) + 64);
    dsc = dalloc<Scalars>(1);
    HIP_CHECK(hipHostMalloc((void**)&hsc, sizeof(Scalars), hipHostMallocDefault));
    HIP_CHECK(hipHostMalloc((void**)&hpart, ((size_t)nblocks(n)*(1 + 2*3) + 64)*sizeof(double), hipHostMallocDefault));
    std::memset(hsc, 0, sizeof(Scalars));
    hsc->lastDt = cfg.last_dt;
    HIP_CHECK(hipMemcpy(dsc, hsc, sizeof(Scalars), hipMemcpyHostToDevice));
    ensureNbrCapacity(64);
  }

  void ensureNbrCapacity(int cap) {
    if (cap <= f.cap) return;
    uint32_t* p = nullptr;
    HIP_CHECK(hipMalloc((void**)&p, (size_t)cap*ld*sizeof(uint32_t)));
    if (f.nbr) {
      auto it = std::find(allocs.begin(), allocs.end(), (void*)f.nbr);
      if (it != allocs.end()) allocs.erase(it);
      HIP_CHECK(hipFree(f.nbr));
    }
    allocs.push_back(p);
    f.nbr = p;
    f.cap = cap;
  }

  // ---------------------------------------------------------------- state I/O
  void setState(const double* x, const double* v, const double* m, const double* rho, const double* eps, const double* H) {
    if (!x || !v || !m || !rho || !eps || !H) raise(SPX_ERR_ARG, "null state pointer");
    std::vector<double> buf((size_t)n);
    auto up = [&](double* dst) { HIP_CHECK(hipMemcpy(dst, buf.data(), n*sizeof(double), hipMemcpyHostToDevice)); };
    for (int a = 0; a < D; ++a) {
      for (int64_t i = 0; i < n; ++i) buf[i] = x[i*D + a];
      up(f.x[a]);
      for (int64_t i = 0; i < n; ++i) buf[i] = v[i*D + a];
      up(f.v[a]);
    }
    for (int64_t i = 0; i < n; ++i) {
      const double* Hi = H + i*D*D;
      const double h = Hi[0];
      for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b) {
          const double want = a == b ? h : 0.0;
          if (Hi[a*D + b] != want)
            raise(SPX_ERR_UNSUPPORTED, "H must be isotropic (h^-1 I); ASPH smoothing tensors are not implemented");
        }
      if (!(h > 0.0) || !std::isfinite(h)) raise(SPX_ERR_ARG, "H must be positive definite");
      if (!(m[i] > 0.0) || !(rho[i] > 0.0)) raise(SPX_ERR_ARG, "mass and mass density must be positive");
      buf[i] = h;
    }
    up(f.hinv);
    std::copy(m, m + n, buf.begin()); up(f.m);
    std::copy(rho, rho + n, buf.begin()); up(f.rho);
    std::copy(eps, eps + n, buf.begin()); up(f.eps);
    for (int64_t i = 0; i < n; ++i) buf[i] = 1.0;
    up(f.omega);
    std::vector<int64_t> ids(n);
    for (int64_t i = 0; i < n; ++i) ids[i] = i;
    HIP_CHECK(hipMemcpy(f.oid, ids.data(), n*sizeof(int64_t), hipMemcpyHostToDevice));
    // derivatives start at zero (SPHBase::initializeProblemStartup)
    zeroDerivs(true);
    HIP_CHECK(hipStreamSynchronize(stream));
    haveState = true;
    initialized = false;
    connValid = false;
    wrapPending = false;
    t = 0.0;
    lastDt = cfg.last_dt;
    cycle = 0;
    hsc->t = 0.0;
    hsc->lastDt = lastDt;
    hsc->cycle = 0;
    hsc->gravDt = 0.0;
    hsc->dtMinAcc = 0.0;
    HIP_CHECK(hipMemcpy(dsc, hsc, sizeof(Scalars), hipMemcpyHostToDevice));
  }

  void zeroDerivs(bool carriedToo) {
    auto z = [&](double* p, size_t cnt) { if (p) HIP_CHECK(hipMemsetAsync(p, 0, cnt*sizeof(double), stream)); };
    for (int a = 0; a < D; ++a) { z(f.DxDt[a], ld); z(f.DvDt[a], ld); z(f.DvDtG[a], ld); z(f.gradRho[a], ld); z(f.m1[a], ld); }
    for (double* p : {f.DepsDt, f.DrhoDt, f.DHDt, f.maxQ, f.Hideal, f.m0, f.rhoSum, f.norm, f.effQ}) z(p, ld);
    z(f.DvDx, (size_t)D*D*ld);
    z(f.Mcorr, (size_t)D*D*ld);
    (void)carriedToo;
  }

  void download(double* dst, const double* src) {
    HIP_CHECK(hipStreamSynchronize(stream));
    HIP_CHECK(hipMemcpy(dst, src, n*sizeof(double), hipMemcpyDeviceToHost));
  }

  int64_t getField(const std::string& name, double* out, int64_t capacity) {
    if (!haveState) raise(SPX_ERR_STATE, "no state uploaded");
    if (wrapPending) applyPendingBoundaries();
    HIP_CHECK(hipStreamSynchronize(stream));
    std::vector<int64_t> oid(n);
    HIP_CHECK(hipMemcpy(oid.data(), f.oid, n*sizeof(int64_t), hipMemcpyDeviceToHost));
    std::vector<double> tmp(n);
    auto scalar = [&](const double* src) -> int64_t {
      if (capacity < n) raise(SPX_ERR_ARG, "output capacity too small");
      download(tmp.data(), src);
      for (int64_t p = 0; p < n; ++p) out[oid[p]] = tmp[p];
      return n;
    };
    auto vec = [&](double* const* src) -> int64_t {
      if (capacity < n*D) raise(SPX_ERR_ARG, "output capacity too small");
      for (int a = 0; a < D; ++a) {
        download(tmp.data(), src[a]);
        for (int64_t p = 0; p < n; ++p) out[oid[p]*D + a] = tmp[p];
      }
      return n*D;
    };
    auto ten = [&](const double* src) -> int64_t {
      if (capacity < n*D*D) raise(SPX_ERR_ARG, "output capacity too small");
      for (int c = 0; c < D*D; ++c) {
        download(tmp.data(), src + (size_t)c*ld);
        for (int64_t p = 0; p < n; ++p) out[oid[p]*D*D + c] = tmp[p];
      }
      return n*D*D;
    };
    auto iso = [&](const double* src) -> int64_t {
      if (capacity < n*D*D) raise(SPX_ERR_ARG, "output capacity too small");
      download(tmp.data(), src);
      for (int64_t p = 0; p < n; ++p)
        for (int a = 0; a < D; ++a)
          for (int b = 0; b < D; ++b) out[oid[p]*D*D + a*D + b] = a == b ? tmp[p] : 0.0;
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
    if (name == "grad h corrections") return scalar(f.omega);
    if (name == "delta velocity hydro") return vec(f.DvDt);
    if (name == "delta velocity") return vec(f.DvDtG);
    if (name == "delta position") return vec(f.DxDt);
    if (name == "delta specific thermal energy") return scalar(f.DepsDt);
    if (name == "delta mass density") return scalar(f.DrhoDt);
    if (name == "velocity gradient") return ten(f.DvDx);
    if (name == "delta H") return iso(f.DHDt);
    if (name == "new H") return iso(f.Hideal);
    if (name == "max viscous pressure") return scalar(f.maxQ);
    if (name == "effective viscous pressure") return scalar(f.effQ);
    if (name == "mass zeroth moment") return scalar(f.m0);
    if (name == "mass first moment") return vec(f.m1);
    if (name == "M SPH correction") return ten(f.Mcorr);
    if (name == "mass density gradient") return vec(f.gradRho);
    if (name == "normalization") return scalar(f.norm);
    if (name == "new mass density") return scalar(f.rhoSum);
    if (name == "gravitational potential") return scalar(f.pot);
    raise(SPX_ERR_ARG, "unknown field '" + name + "'");
  }

  // ---------------------------------------------------------------- connectivity
  void setupGrid() {
    const int nb = nblocks(n);
    const int nv = 1 + 2*D;
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_extent<DD>, dim3(nb), dim3(kBlock), 0, stream, f, (int)n, part);
    });
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(hpart, part, (size_t)nb*nv*sizeof(double), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    double hmin = 1.7976931348623157e308, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) { lo[a] = 1.7976931348623157e308; hi[a] = -1.7976931348623157e308; }
    for (int b = 0; b < nb; ++b) {
      hmin = std::min(hmin, hpart[b*nv]);
      for (int a = 0; a < D; ++a) { lo[a] = std::min(lo[a], hpart[b*nv + 1 + 2*a]); hi[a] = std::max(hi[a], -hpart[b*nv + 2 + 2*a]); }
    }
    if (!(hmin > 0.0) || !std::isfinite(hmin)) raise(SPX_ERR_NUMERIC, "non-positive or non-finite H on the device");
    const double support = cfg.neighbor_extent/hmin*(1.0 + 1.0e-9);
    int64_t total = 1;
    for (int a = 0; a < D; ++a) {
      const int blo = geo.bc_lo[a], bhi = geo.bc_hi[a];
      const double L = cfg.xmax[a] - cfg.xmin[a];
      int64_t nc;
      double cell, origin;
      if (blo == BC_PERIODIC) {
        nc = (int64_t)std::floor(L/support);
        if (nc < 3) raise(SPX_ERR_UNSUPPORTED, "periodic box must span at least 3 support radii on every axis");
        nc = std::min<int64_t>(nc, 1 << 20);
        cell = L/nc;
        origin = cfg.xmin[a];
      } else if (blo == BC_REFLECTING && bhi == BC_REFLECTING) {
        nc = std::max<int64_t>(1, std::min<int64_t>((int64_t)std::floor(L/support), 1 << 20));
        cell = L/nc;
        origin = cfg.xmin[a];
      } else if (blo == BC_REFLECTING) {
        const double top = std::max(hi[a], cfg.xmin[a]);
        nc = std::max<int64_t>(1, std::min<int64_t>((int64_t)std::floor((top - cfg.xmin[a])/support) + 1, 1 << 20));
        cell = std::max(support, (top - cfg.xmin[a])/nc*(1.0 + 1e-12));
        origin = cfg.xmin[a];
      } else if (bhi == BC_REFLECTING) {
        const double bot = std::min(lo[a], cfg.xmax[a]);
        nc = std::max<int64_t>(1, std::min<int64_t>((int64_t)std::floor((cfg.xmax[a] - bot)/support) + 1, 1 << 20));
        cell = std::max(support, (cfg.xmax[a] - bot)/nc*(1.0 + 1e-12));
        origin = cfg.xmax[a] - nc*cell;
      } else {
        const double span = std::max(0.0, hi[a] - lo[a]);
        nc = std::max<int64_t>(1, std::min<int64_t>((int64_t)std::floor(span/support) + 1, 1 << 20));
        cell = std::max(support, span/nc*(1.0 + 1e-12));
        origin = lo[a];
      }
      geo.nc[a] = (int)nc;
      geo.cell[a] = cell;
      geo.inv_cell[a] = 1.0/cell;
      geo.origin[a] = origin;
      total *= nc;
    }
    if (total > (int64_t)1 << 27) raise(SPX_ERR_CAPACITY, "cell grid too large (support radius tiny relative to the domain)");
    if (total > cellsAlloc) {
      cellStart = dalloc<int>(total);
      cellEnd = dalloc<int>(total);
      cellsAlloc = total;
    }
    HIP_CHECK(hipMemsetAsync(cellStart, 0, total*sizeof(int), stream));
    HIP_CHECK(hipMemsetAsync(cellEnd, 0, total*sizeof(int), stream));
  }

  void buildConnectivity() {
    if (!haveState) raise(SPX_ERR_STATE, "no state uploaded");
    setupGrid();
    const int nb = nblocks(n);
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_keys<DD>, dim3(nb), dim3(kBlock), 0, stream, f, geo, (int)n, wrapPending ? 1 : 0, keys, vals);
    });
    HIP_CHECK(hipGetLastError());
    wrapPending = false;
    int bits = 0;
    for (int a = 0; a < D; ++a) { int b = 0; while ((1 << b) < geo.nc[a]) ++b; bits = std::max(bits, b); }
    const int endBit = std::max(1, bits*D);
    size_t tmp = sortTmpBytes;
    HIP_CHECK(rocprim::radix_sort_pairs(sortTmp, tmp, keys, keysAlt, vals, valsAlt, (size_t)n, 0, endBit, stream));
    // reorder the carried arrays into cell order
    const bool reorder = cfg.reorder_interval <= 1 || stepsSinceReorder + 1 >= cfg.reorder_interval || !initialized;
    const uint32_t* sorted = valsAlt;
    if (reorder) {
      auto src = carried(f);
      auto dst = carried(alt);
      GatherSet gs{};
      gs.count = 0;
      for (size_t k = 0; k < src.size(); ++k) { gs.src[gs.count] = *src[k]; gs.dst[gs.count] = *dst[k]; ++gs.count; }
      for (int c = 0; c < D*D; ++c) { gs.src[gs.count] = f.DvDx + (size_t)c*ld; gs.dst[gs.count] = alt.DvDx + (size_t)c*ld; ++gs.count; }
      hipLaunchKernelGGL(k_gather, dim3(nb), dim3(kBlock), 0, stream, gs, valsAlt, (int)n);
      hipLaunchKernelGGL(k_gather_i64, dim3(nb), dim3(kBlock), 0, stream, f.oid, oidAlt, valsAlt, (int)n);
      HIP_CHECK(hipGetLastError());
      for (size_t k = 0; k < src.size(); ++k) std::swap(*src[k], *dst[k]);
      std::swap(f.DvDx, alt.DvDx);
      std::swap(f.oid, oidAlt);
      sorted = nullptr;
      stepsSinceReorder = 0;
    } else {
      ++stepsSinceReorder;
    }
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_cell_bounds<DD>, dim3(nb), dim3(kBlock), 0, stream, keysAlt, (int)n, geo, cellStart, cellEnd);
    });
    HIP_CHECK(hipGetLastError());
    for (int attempt = 0; attempt < 4; ++attempt) {
      hsc->overflow = 0;
      HIP_CHECK(hipMemcpyAsync(&dsc->overflow, &hsc->overflow, sizeof(int), hipMemcpyHostToDevice, stream));
      launchD([&](auto d) {
        constexpr int DD = decltype(d)::value;
        hipLaunchKernelGGL(k_neighbors<DD>, dim3(nb), dim3(kBlock), 0, stream, f, geo, ph, (int)n, cellStart, cellEnd,
                           sorted, dsc);
      });
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemcpyAsync(&hsc->overflow, &dsc->overflow, sizeof(int), hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      if (hsc->overflow == 0) break;
      const int need = hsc->overflow;
      if (need > 4096) raise(SPX_ERR_CAPACITY, "more than 4096 neighbours for one node");
      ensureNbrCapacity(((need + need/4) + 7)/8*8);
    }
    connValid = true;
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

  void applyPendingBoundaries() {
    // enforce boundaries without a neighbour rebuild: key kernel on a scratch grid
    const int nb = nblocks(n);
    Geo g = geo;
    for (int a = 0; a < 3; ++a) { g.nc[a] = 1; g.inv_cell[a] = 0.0; g.origin[a] = 0.0; }
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_keys<DD>, dim3(nb), dim3(kBlock), 0, stream, f, g, (int)n, 1, keys, vals);
    });
    HIP_CHECK(hipGetLastError());
    wrapPending = false;
  }

  // ---------------------------------------------------------------- package passes
  void copyStateToMid() {
    auto cp = [&](double* d, const double* s) { HIP_CHECK(hipMemcpyAsync(d, s, n*sizeof(double), hipMemcpyDeviceToDevice, stream)); };
    for (int a = 0; a < D; ++a) { cp(f.xm[a], f.x[a]); cp(f.vm[a], f.v[a]); }
    cp(f.rhom, f.rho); cp(f.epsm, f.eps); cp(f.hinvm, f.hinv); cp(f.Pm, f.P); cp(f.csm, f.cs); cp(f.omegam, f.omega);
  }
  void eos() { hipLaunchKernelGGL(k_eos, dim3(nblocks(n)), dim3(kBlock), 0, stream, f, ph, (int)n); HIP_CHECK(hipGetLastError()); }
  void omegaPass(bool mid) {
    if (!cfg.gradh_correction) return;
    double** xs = mid ? xsMid : xsState;
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_omega<DD>, dim3(nblocks(n)), dim3(kBlock), 0, stream, f, geo, dtab, (int)n, xs,
                         mid ? f.hinvm : f.hinv, mid ? f.omegam : f.omega);
    });
    HIP_CHECK(hipGetLastError());
  }
  double** xsMid = nullptr;    // device arrays of position pointers
  double** xsState = nullptr;
  void refreshPosPtrs() {
    if (!xsMid) { xsMid = dalloc<double*>(3); xsState = dalloc<double*>(3); }
    double* hm[3] = {f.xm[0], f.xm[1], f.xm[2]};
    double* hs[3] = {f.x[0], f.x[1], f.x[2]};
    HIP_CHECK(hipMemcpyAsync(xsMid, hm, sizeof(hm), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(xsState, hs, sizeof(hs), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  }
  void densityDt(bool dens, bool dtv) {
    const int nb = nblocks(n);
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      if (dens && dtv) hipLaunchKernelGGL((k_density_dt<DD, true, true>), dim3(nb), dim3(kBlock), 0, stream, f, geo, ph, dtab, (int)n, part);
      else if (dens) hipLaunchKernelGGL((k_density_dt<DD, true, false>), dim3(nb), dim3(kBlock), 0, stream, f, geo, ph, dtab, (int)n, part);
      else hipLaunchKernelGGL((k_density_dt<DD, false, true>), dim3(nb), dim3(kBlock), 0, stream, f, geo, ph, dtab, (int)n, part);
    });
    HIP_CHECK(hipGetLastError());
  }
  void sphPass(bool full) {
    const int nb = nblocks(n);
    hipLaunchKernelGGL(k_prho, dim3(nb), dim3(kBlock), 0, stream, f, (int)n);
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      if (full) hipLaunchKernelGGL((k_sph<DD, true>), dim3(nb), dim3(kBlock), 0, stream, f, geo, ph, dtab, (int)n);
      else hipLaunchKernelGGL((k_sph<DD, false>), dim3(nb), dim3(kBlock), 0, stream, f, geo, ph, dtab, (int)n);
    });
    HIP_CHECK(hipGetLastError());
  }
  // TreeGravity::initialize (tree on the given positions) + evaluateDerivatives + dt vote
  void gravityPass(bool mid) {
    if (!cfg.gravity) return;
    double* xs[3] = {mid ? f.xm[0] : f.x[0], mid ? f.xm[1] : f.x[1], mid ? f.xm[2] : f.x[2]};
    grav.build(xs, f.m, (int)n);
    grav.evaluate(f.DvDtG, f.pot, (int)n, cfg.G, cfg.softening, cfg.opening, dsc, cfg.ftimestep);
  }

  void mark(const char* name) {
    if (!timing) return;
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipEventRecord(e, stream));
    evs.push_back(e);
    tnames.push_back(name);
  }
  void collectTimes() {
    if (!timing || evs.empty()) return;
    HIP_CHECK(hipEventSynchronize(evs.back()));
    std::vector<std::string> names;
    std::vector<double> ms;
    for (size_t k = 1; k < evs.size(); ++k) {
      float v = 0.f;
      HIP_CHECK(hipEventElapsedTime(&v, evs[k - 1], evs[k]));
      const std::string& nm = tnames[k];
      auto it = std::find(names.begin(), names.end(), nm);
      if (it == names.end()) { names.push_back(nm); ms.push_back(v); }
      else ms[it - names.begin()] += v;
    }
    for (auto e : evs) (void)hipEventDestroy(e);
    evs.clear();
    tnames.clear();
    this->tnames = names;
    this->tms = ms;
  }

  // ---------------------------------------------------------------- lifecycle
  void initialize() {
    if (!haveState) raise(SPX_ERR_STATE, "no state uploaded");
    initialized = false;
    buildConnectivity();
    initialized = true;
    refreshPosPtrs();
    zeroDerivs(true);
    eos();
    omegaPass(false);
    if (cfg.gravity) gravityPass(false);
    HIP_CHECK(hipStreamSynchronize(stream));
  }

  double step(double maxTime) {
    if (!initialized) raise(SPX_ERR_STATE, "spx_run_initialize must be called before stepping");
    if (timing) { tnames.clear(); for (auto e : evs) (void)hipEventDestroy(e); evs.clear(); mark("start"); }
    // Integrator::step -> setGhostNodes / connectivity
    buildConnectivity();
    refreshPosPtrs();
    mark("connectivity");
    // preStepInitialize: RigorousSumDensity, fused with the dt vote
    densityDt(true, true);
    mark("density_dt");
    // TreeGravity::initialize builds the tree on the start-of-step positions; the
    // vote uses the previous evaluation's dtMinAcc.
    hipLaunchKernelGGL(k_select_dt, dim3(1), dim3(kBlock), 0, stream, part, nblocks(n), dsc, ph,
                       cfg.dt_min, cfg.dt_max, maxTime, cfg.dt_growth, cfg.gravity);
    HIP_CHECK(hipGetLastError());
    // mid-step increments with the previous derivatives
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL(k_midstep<DD>, dim3(nblocks(n)), dim3(kBlock), 0, stream, f, ph, dsc, (int)n);
    });
    HIP_CHECK(hipGetLastError());
    mark("update");
    omegaPass(true);
    mark("omega");
    // derivatives at the midpoint
    if (cfg.gravity) {
      gravityPass(true);
      mark("gravity");
    }
    sphPass(false);
    mark("sph");
    // full step with compatible energy
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL((k_full_update<DD, false>), dim3(nblocks(n)), dim3(kBlock), 0, stream, f, geo, ph, dtab, dsc, 0.0, (int)n);
    });
    HIP_CHECK(hipGetLastError());
    for (int a = 0; a < D; ++a) std::swap(f.v[a], f.v1[a]);
    mark("energy_update");
    omegaPass(false);
    mark("omega");
    wrapPending = true;
    // read dt back
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, stream, dsc);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(hsc, dsc, sizeof(Scalars), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    collectTimes();
    t = hsc->t;
    lastDt = hsc->lastDt;
    cycle = hsc->cycle;
    return hsc->lastDt;
  }

  void evaluateDerivatives() {
    if (!connValid) raise(SPX_ERR_STATE, "no connectivity: call spx_run_update_connectivity or spx_run_initialize");
    copyStateToMid();
    zeroDerivs(false);
    if (cfg.gravity) gravityPass(true);
    sphPass(true);
    HIP_CHECK(hipStreamSynchronize(stream));
  }

  void hydroDtVote(double* hydro, double* gravity) {
    if (!connValid) raise(SPX_ERR_STATE, "no connectivity");
    densityDt(false, true);
    hipLaunchKernelGGL(k_select_dt, dim3(1), dim3(kBlock), 0, stream, part, nblocks(n), dsc, ph, 0.0,
                       1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308, cfg.gravity);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(hsc, dsc, sizeof(Scalars), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    *hydro = hsc->hydroDt;
    *gravity = cfg.gravity ? hsc->gravDt : 0.0;
  }

  void compatibleEnergy(double multiplier) {
    if (!connValid) raise(SPX_ERR_STATE, "no connectivity");
    launchD([&](auto d) {
      constexpr int DD = decltype(d)::value;
      hipLaunchKernelGGL((k_full_update<DD, true>), dim3(nblocks(n)), dim3(kBlock), 0, stream, f, geo, ph, dtab, dsc, multiplier, (int)n);
    });
    HIP_CHECK(hipGetLastError());
    eos();
    HIP_CHECK(hipStreamSynchronize(stream));
  }

  void neighbors(int64_t* offsets, int64_t* nbr, int32_t* img, int64_t capacity, int32_t* countsOnly) {
    if (!connValid) raise(SPX_ERR_STATE, "no connectivity");
    HIP_CHECK(hipStreamSynchronize(stream));
    std::vector<int32_t> cnt(n);
    std::vector<int64_t> oid(n);
    HIP_CHECK(hipMemcpy(cnt.data(), f.cnt, n*sizeof(int32_t), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(oid.data(), f.oid, n*sizeof(int64_t), hipMemcpyDeviceToHost));
    if (countsOnly) {
      for (int64_t p = 0; p < n; ++p) countsOnly[oid[p]] = cnt[p];
      return;
    }
    std::vector<uint32_t> ell((size_t)f.cap*ld);
    HIP_CHECK(hipMemcpy(ell.data(), f.nbr, ell.size()*sizeof(uint32_t), hipMemcpyDeviceToHost));
    std::vector<int64_t> off(n + 1, 0);
    std::vector<int64_t> slotOf(n);
    for (int64_t p = 0; p < n; ++p) { off[oid[p] + 1] = cnt[p]; slotOf[oid[p]] = p; }
    for (int64_t i = 0; i < n; ++i) off[i + 1] += off[i];
    if (capacity < off[n]) raise(SPX_ERR_ARG, "neighbour capacity too small");
    for (int64_t i = 0; i < n; ++i) {
      const int64_t p = slotOf[i];
      std::vector<std::pair<int64_t, int32_t>> l;
      for (int k = 0; k < cnt[p]; ++k) {
        const uint32_t e = ell[(size_t)k*ld + p];
        l.push_back({oid[e & kIdxMask], (int32_t)(e >> kImgShift)});
      }
      std::sort(l.begin(), l.end());
      for (size_t k = 0; k < l.size(); ++k) { nbr[off[i] + k] = l[k].first; img[off[i] + k] = l[k].second; }
      offsets[i] = off[i];
    }
    offsets[n] = off[n];
  }
  int64_t totalNeighbors() {
    std::vector<int32_t> cnt(n);
    HIP_CHECK(hipStreamSynchronize(stream));
    HIP_CHECK(hipMemcpy(cnt.data(), f.cnt, n*sizeof(int32_t), hipMemcpyDeviceToHost));
    int64_t s = 0;
    for (auto c : cnt) s += c;
    return s;
  }
  int64_t numPairs() {
    if (!connValid) raise(SPX_ERR_STATE, "no connectivity");
    if (npairsCache >= 0) return npairsCache;
    // unique pairs: internal pairs counted once, image pairs once per (node, image) entry
    std::vector<int32_t> cnt(n);
    HIP_CHECK(hipStreamSynchronize(stream));
    HIP_CHECK(hipMemcpy(cnt.data(), f.cnt, n*sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<uint32_t> ell((size_t)f.cap*ld);
    HIP_CHECK(hipMemcpy(ell.data(), f.nbr, ell.size()*sizeof(uint32_t), hipMemcpyDeviceToHost));
    int64_t np = 0;
    for (int64_t p = 0; p < n; ++p)
      for (int k = 0; k < cnt[p]; ++k) {
        const uint32_t e = ell[(size_t)k*ld + p];
        const int64_t j = e & kIdxMask;
        if ((e >> kImgShift) != 0 || j > p) ++np;
      }
    npairsCache = np;
    return np;
  }

  void conservation(double* out) {
    if (!haveState) raise(SPX_ERR_STATE, "no state");
    std::vector<double> m(n), e(n), v[3], pot(n);
    download(m.data(), f.m);
    download(e.data(), f.eps);
    download(pot.data(), f.pot);
    for (int a = 0; a < D; ++a) { v[a].resize(n); download(v[a].data(), f.v[a]); }
    double M = 0, P[3] = {0, 0, 0}, KE = 0, TE = 0, PE = 0;
    for (int64_t i = 0; i < n; ++i) {
      M += m[i];
      double v2 = 0.0;
      for (int a = 0; a < D; ++a) { P[a] += m[i]*v[a][i]; v2 += v[a][i]*v[a][i]; }
      KE += 0.5*m[i]*v2;
      TE += m[i]*e[i];
      PE += 0.5*m[i]*pot[i];
    }
    out[0] = M; out[1] = P[0]; out[2] = P[1]; out[3] = P[2]; out[4] = KE; out[5] = TE; out[6] = cfg.gravity ? PE : 0.0;
  }
};

}  // namespace spx

struct spx_run {
  std::unique_ptr<spx::Run> impl;
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
    t.build(cfg->kernel, cfg->dim, (unsigned)cfg->kernel_points, cfg->min_nperh, cfg->max_nperh);
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
    t.build(cfg->kernel, cfg->dim, (unsigned)cfg->kernel_points, cfg->min_nperh, cfg->max_nperh);
    *ncoef = (int64_t)t.W.coef.size();
    if (wc) std::memcpy(wc, t.W.coef.data(), t.W.coef.size()*sizeof(double));
    if (gwc) std::memcpy(gwc, t.gW.coef.data(), t.gW.coef.size()*sizeof(double));
    if (xstep) *xstep = t.W.xstep;
  });
}

int spx_run_create(const spx_config* cfg, int64_t n, int32_t device, spx_run** out) {
  return guarded([&] {
    if (!cfg || !out) spx::raise(SPX_ERR_ARG, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) spx::raise(SPX_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) spx::raise(SPX_ERR_ARG, "device index out of range");
    auto r = new spx_run;
    try {
      r->impl = std::make_unique<spx::Run>(*cfg, n, device);
    } catch (...) {
      delete r;
      throw;
    }
    *out = r;
  });
}

void spx_run_destroy(spx_run* run) { delete run; }

int spx_run_set_state(spx_run* run, const double* x, const double* v, const double* m, const double* rho,
                      const double* eps, const double* H) {
  RUN_OR_FAIL(run);
  return guarded([&] { run->impl->setState(x, v, m, rho, eps, H); });
}

int spx_run_get_field(spx_run* run, const char* name, double* out, int64_t capacity, int64_t* count) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!name || !out) spx::raise(SPX_ERR_ARG, "null argument");
    const int64_t c = run->impl->getField(name, out, capacity);
    if (count) *count = c;
  });
}

int spx_run_initialize(spx_run* run) {
  RUN_OR_FAIL(run);
  return guarded([&] { run->impl->initialize(); });
}

int spx_run_step(spx_run* run, double max_time, double* dt_out) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    const double dt = run->impl->step(max_time);
    if (dt_out) *dt_out = dt;
  });
}

int spx_run_advance(spx_run* run, double max_time, int64_t nsteps, int64_t* steps_done, double* t_out) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    int64_t k = 0;
    auto& R = *run->impl;
    while (k < nsteps && R.t < max_time*(1.0 - 1.0e-15)) { R.step(max_time); ++k; }
    if (steps_done) *steps_done = k;
    if (t_out) *t_out = R.t;
  });
}

int spx_run_time(spx_run* run, double* t, int64_t* cycle, double* last_dt) {
  RUN_OR_FAIL(run);
  if (t) *t = run->impl->t;
  if (cycle) *cycle = run->impl->cycle;
  if (last_dt) *last_dt = run->impl->lastDt;
  return SPX_OK;
}

int spx_run_update_connectivity(spx_run* run) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    run->impl->buildConnectivity();
    run->impl->refreshPosPtrs();
  });
}

int spx_run_neighbor_counts(spx_run* run, int32_t* counts) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!counts) spx::raise(SPX_ERR_ARG, "null argument");
    run->impl->neighbors(nullptr, nullptr, nullptr, 0, counts);
  });
}

int spx_run_neighbors(spx_run* run, int64_t* offsets, int64_t* nbr, int32_t* img, int64_t capacity) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!offsets || !nbr || !img) spx::raise(SPX_ERR_ARG, "null argument");
    run->impl->neighbors(offsets, nbr, img, capacity, nullptr);
  });
}

int spx_run_sum_density(spx_run* run) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!run->impl->connValid) spx::raise(SPX_ERR_STATE, "no connectivity");
    run->impl->densityDt(true, false);
    HIP_CHECK(hipStreamSynchronize(run->impl->stream));
  });
}

int spx_run_update_eos(spx_run* run) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    run->impl->eos();
    HIP_CHECK(hipStreamSynchronize(run->impl->stream));
  });
}

int spx_run_post_state_update(spx_run* run) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!run->impl->connValid) spx::raise(SPX_ERR_STATE, "no connectivity");
    run->impl->omegaPass(false);
    HIP_CHECK(hipStreamSynchronize(run->impl->stream));
  });
}

int spx_run_evaluate_derivatives(spx_run* run, double, double) {
  RUN_OR_FAIL(run);
  return guarded([&] { run->impl->evaluateDerivatives(); });
}

int spx_run_dt(spx_run* run, double* hydro_dt, double* gravity_dt) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    double h = 0, g = 0;
    run->impl->hydroDtVote(&h, &g);
    if (hydro_dt) *hydro_dt = h;
    if (gravity_dt) *gravity_dt = g;
  });
}

int spx_run_compatible_energy(spx_run* run, double multiplier) {
  RUN_OR_FAIL(run);
  return guarded([&] { run->impl->compatibleEnergy(multiplier); });
}

int spx_run_conservation(spx_run* run, double* out7) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!out7) spx::raise(SPX_ERR_ARG, "null argument");
    run->impl->conservation(out7);
  });
}

int spx_run_kernel_times(spx_run* run, double* ms, int32_t max_entries, int32_t* n_entries, char* names, int64_t names_cap) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    auto& R = *run->impl;
    if (max_entries < 0) { R.timing = true; return; }   // enable timing
    const int32_t k = (int32_t)std::min<size_t>(R.tms.size(), (size_t)max_entries);
    std::string all;
    for (int32_t i = 0; i < k; ++i) { if (ms) ms[i] = R.tms[i]; all += (i ? "," : "") + R.tnames[i]; }
    if (n_entries) *n_entries = k;
    if (names && names_cap > 0) {
      const size_t c = std::min<size_t>(all.size(), (size_t)names_cap - 1);
      std::memcpy(names, all.data(), c);
      names[c] = 0;
    }
  });
}

int spx_run_num_pairs(spx_run* run, int64_t* npairs) {
  RUN_OR_FAIL(run);
  return guarded([&] {
    if (!npairs) spx::raise(SPX_ERR_ARG, "null argument");
    *npairs = run->impl->numPairs();
  });
}

}  // extern "C"
