// spx_device.hpp -- device-side types and helpers shared by the engine's kernels.
//
// Data layout in HBM: structure of arrays, one double array per scalar component,
// particles in cell-Morton order (slot p holds original node oid[p]).  H is
// isotropic (H = hinv * I, as SPH + IdealH keep it), so one double per node.
// Neighbour lists are a blocked ELLPACK: entries come in chunks of 8 per slot, chunk c of
// slot i at nbr[(c*ld + i)*8 .. +8) (ld >= n), so a lane writes its list 32 bytes at a
// time and a wavefront reading chunk c reads 2 KB contiguously.  Each entry is a uint32
// holding the neighbour slot in the low 26 bits and a 6-bit image code (2 bits per axis:
// 0 none, 1 beyond the low plane, 2 beyond the high plane).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spx {

constexpr int kImgShift = 26;
constexpr uint32_t kIdxMask = (1u << kImgShift) - 1u;
constexpr int kBlock = 256;
constexpr int kEllChunk = 8;

// Position of neighbour entry k of slot i in the blocked ELLPACK array.
__host__ __device__ __forceinline__ size_t ell_at(int k, int i, size_t ld) {
  return (((size_t)(k >> 3)*ld + (size_t)i) << 3) + (size_t)(k & 7);
}

// Neighbour entries are consumed in groups of G (4 or 8): one or two 16-byte loads of the
// blocked list, then every neighbour read of the group is issued before any is used, so a
// wavefront waits on memory once per group instead of twice per neighbour.  Positions past the end of the list
// (k0 + u >= nn) become the lane's own slot with no image, and are skipped by the caller.
// The filtering sum-density + dt pass: 4 in flight (137 VGPRs, 3 waves/SIMD) 1.28 ms per step on
// FP64 Taylor-Green 128^3; 8 in flight (219 VGPRs, 2 waves) 1.32 ms.
constexpr int kGroup = 4;
// The compatible-energy update does little arithmetic per neighbour; it keeps 8 in flight.
// FP64 Taylor-Green 128^3 per step: 8 in flight (184 VGPRs, 2 waves/SIMD) 0.96 ms; 4 in flight
// (102 VGPRs, 4 waves) 1.16 ms; 8 capped at 3 waves (167 VGPRs, no spills) 1.04 ms.
constexpr int kGroupUpdate = 8;
// The grad-h passes read 32 bytes per neighbour and evaluate one kernel lookup: 8 in flight
// (88 VGPRs, 5 waves/SIMD) measured 0.92 -> 0.906 ms FP64 per step on Taylor-Green 128^3.
constexpr int kGroupOmega = 8;
template <int G>
__device__ __forceinline__ void ell_group(const uint32_t* nbr, size_t ld, int k0, int i, int nn, int* js, int* codes) {
  uint32_t e[G];
  const uint32_t* p = nbr + ell_at(k0, i, ld);
  if (G == 8) {
    const uint4 q = *reinterpret_cast<const uint4*>(p), r = *reinterpret_cast<const uint4*>(p + 4);
    e[0] = q.x; e[1 % G] = q.y; e[2 % G] = q.z; e[3 % G] = q.w;
    e[4 % G] = r.x; e[5 % G] = r.y; e[6 % G] = r.z; e[7 % G] = r.w;
  } else if (G == 4) {
    const uint4 q = *reinterpret_cast<const uint4*>(p);
    e[0] = q.x; e[1 % G] = q.y; e[2 % G] = q.z; e[3 % G] = q.w;
  } else if (G == 2) {
    const uint2 q = *reinterpret_cast<const uint2*>(p);
    e[0] = q.x; e[1 % G] = q.y;
  } else {
    e[0] = p[0];
  }
#pragma unroll
  for (int u = 0; u < G; ++u) {
    const bool ok = k0 + u < nn;
    js[u] = ok ? (int)(e[u] & kIdxMask) : i;
    codes[u] = ok ? (int)(e[u] >> kImgShift) : 0;
  }
}

// The command processor hands workgroup b to XCD b % 8, and each XCD has its own L2.
// Pair kernels map blocks so that each XCD works through one contiguous range of
// (Morton-ordered) particles and keeps its neighbours' data in its own L2.  Grids of
// such kernels are launched as a multiple of kXcd blocks (xcd_grid); any other grid
// keeps the identity mapping.
constexpr int kXcd = 8;
__device__ __forceinline__ int xcd_block() {
  const int nb = (int)gridDim.x;
  if (nb % kXcd) return (int)blockIdx.x;
  return ((int)blockIdx.x % kXcd)*(nb/kXcd) + (int)blockIdx.x/kXcd;
}
static inline unsigned xcd_grid(long long n, int block) {
  const long long nb = (n + block - 1)/block;
  return (unsigned)((nb + kXcd - 1)/kXcd*kXcd);
}

enum Bc { BC_NONE = 0, BC_PERIODIC = 1, BC_REFLECTING = 2 };

// Life cycle of a run slot in a batch.
enum RunState { RUN_FREE = 0, RUN_ACTIVE = 1, RUN_DONE = 2, RUN_REMOVED = 3 };

// Tabulated kernel: per interval 6 coefficients (W: c0 c1 c2, dW: c0 c1 c2),
// plus the Wsum -> nPerh cubic Hermite table (N values then N knot gradients).
struct DevTable {
  const double* coef;
  const float* coef32;     // the same coefficients rounded to float (mixed-precision pair math)
  const double* herm;
  int n1;
  int hn;
  double inv_xstep, extent, W0;
  double hxmin, hxmax, hxstep;
};

// Boundary planes, shared by every run of a batch.
struct Geo {
  int bc_lo[3], bc_hi[3];
  double xmin[3], xmax[3];
  // bit a: axis a periodic; bit 2a + side - 1 (side 1 low, 2 high): that side reflecting
  int permask, reflmask;
};

// Cell grid of one run for the current step.  Cells of run r occupy the global
// cell-index range [cellBase, cellBase + nc0*nc1*nc2).
// Cell levels of a run whose h varies (open boundaries): level-l cells are 2^(levels-1-l)
// finest cells per axis, l < kCellLevels.
constexpr int kCellLevels = 8;
struct RunGrid {
  double origin[3];
  double inv_cell[3];  // of the finest level when levels > 1
  double hinvMin;      // smallest h^-1 of the run's nodes: bounds every node's support radius
  double ext2c;        // candidate-list criterion: ext^2 (1 + skin)^2
  double support;      // the largest widened support radius
  int nc[3];           // cells per axis (of the finest level when levels > 1)
  int cellBase;        // first cell of the run's table (of level 0 when levels > 1)
  int levels;          // > 1: cell levels, level l at least support/2^(l+2) wide
  int levelBase[kCellLevels];   // first cell of each level's table
};

// Physics constants of the run.
struct Phys {
  double gamma, g1, Cl, Cq, cfl, nperh, ext, ext2, hmin, hmax, rho_min, rho_max;
  double hminInv, hmaxInv;
  double etaCrit, etaFold;   // LimitedMonaghanGingoldViscosity: etaCritFrac/nPerh, etaFoldFrac/nPerh
  int correctVG, useVmagDt, gravity;
  // the Limited Q's gradient on image partners: 1 the control node's tensor image (boundaries
  // applied after the Q's copy), 0 zero (the integrator's midpoint with grad-h off)
  int qImages;
};

struct Fields {
  // state (start of step, then end of step)
  double *x[3], *v[3], *m, *rho, *eps, *hinv, *P, *cs, *omega;
  // midpoint state
  double *xm[3], *vm[3], *rhom, *epsm, *hinvm, *Pm, *csm, *omegam, *prhom;
  // end-of-step velocity (swapped with v after the full update)
  double *v1[3];
  // derivatives carried across steps
  double *DxDt[3], *DvDt[3], *DvDtG[3], *DepsDt, *DrhoDt, *DHDt, *DvDx, *maxQ;
  // the Limited Q's copy of DvDx from the previous derivative evaluation (D*D planes; null
  // unless that Q is selected)
  double *DvDxQ;
  // derivatives consumed within a step
  double *Hideal, *m0;
  // evaluateDerivatives outputs not consumed by the integration
  double *rhoSum, *norm, *effQ, *gradRho[3], *m1[3], *Mcorr;
  double *ns;        // nodeScale = 1/(lambda_max(H) nPerh) of the connectivity state
  // Packed per-node records, so that a pair pass reads a neighbour with a few 16-byte loads
  // from one line per record instead of one load per array.  Each record array holds only
  // fields its readers use together (a pass that needs x and h^-1 alone reads 32 bytes per
  // node).  Unused axes hold 0.
  //   xh  (4 doubles): x0 x1 x2 h^-1 of the connectivity state (k_pack), then of the
  //                    end-of-step state (k_full_update)
  //   vmr (4):         v0 v1 v2 m of the connectivity state (k_pack)
  //   xhm (4):         x0 x1 x2 h^-1 of the midpoint state (k_midstep)
  //   pkm (8):         v0 v1 v2 m rho cs Prho 1/rho of the midpoint state (midpoint grad-h pass)
  //   pkv (8):         v^{1/2} (3) DepsDt m x (3) for the compatible-energy partition: the
  //                    half-step velocity and DepsDt (SPH pass in the step, else k_v12), m and
  //                    the position the SPH pass evaluated at (SPH pass)
  double *xh, *vmr, *xhm, *pkm, *pkv;
  double *pot;       // gravitational potential of the last TreeGravity evaluation
  // The potential as the integrator's State holds it: TreeGravity enrolls its potential in the
  // State (Gravity/TreeGravity.cc:137), so an integrator's state0 copy and the state.assign(state0)
  // after the midpoint evaluation (Integrator/CheapSynchronousRK2.cc:66-68,107,
  // Integrator/SynchronousRK2.cc:68-69,101) leave it at the start-of-step evaluation (SynchronousRK2)
  // or at its value before the step (CheapSynchronousRK2: the start-up dry run or a standalone
  // evaluation).
  double *potS;
  uint32_t* nbr;
  // Candidate lists (same blocked layout as nbr): every j with |Hi rij| or |Hj rij| within
  // ext (1 + skin) when the run's list was built, from positions xb and h^-1 hb of that time.
  uint32_t* cand;
  int32_t* ccnt;
  double *xb[3], *hb;
  // Pairwise acceleration coefficient of each ELL entry, in the pair arithmetic type: the
  // reference's pairAccelerations of the pair (i, j) is -m_j C r_ij (isotropic H), C of entry k
  // of slot i at pdd[k*ld + i] (the same bits from both sides of the pair).
  void* pdd;
  int32_t* cnt;
  int64_t* oid;      // (run slot << 40) | node index within the run
  int32_t* run;      // run slot of each particle
  int ld, cap, ccap;
};

// Device scalars of one run.
struct Scalars {
  double t, lastDt, dt, hdt;
  double hydroDt, gravDt, dtMinAcc, tEnd;
  long long cycle;
  unsigned long long dtBits;   // running min of the hydro dt vote (order-preserving bits)
  // since the run's candidate list was built (order-preserving bits, 0 = nothing measured):
  // max over nodes of |x - xb| hb/ext, and of hb/h^-1 (growth of the support radius)
  unsigned long long skinD, skinG;
  int active, overflow;
  int nonfinite, pad;
};

// Order-preserving map between doubles and unsigned 64-bit integers (for atomic min/max).
__device__ __forceinline__ unsigned long long ord_bits(double d) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(d);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double ord_value(unsigned long long u) {
  const unsigned long long b = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
  double d;
  __builtin_memcpy(&d, &b, sizeof(d));
  return d;
}

__device__ __forceinline__ double dmul(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double dadd(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ double dsub(double a, double b) { return __dsub_rn(a, b); }

// Ghost coordinate of an image along one axis, written in the reference's
// mapPositionThroughPlanes arithmetic for axis-aligned planes through the origin of
// the other axes (Boundary/mapPositionThroughPlanes.hh:17-25, Geometry/GeomPlane.cc:240-259,
// Utilities/pointDistances.hh:69-74).  Adds only: no contraction possible.
template <int D>
__device__ __forceinline__ double image_coord(int bc, int code, double x, double lo, double hi) {
  if (bc == BC_PERIODIC) {
    if (code == 2) {
      if (D == 3) { const double cl = dadd(x, dsub(hi, x)); return dadd(cl, dsub(x, lo)); }
      return dadd(hi, dsub(x, lo));
    }
    if (D == 3) { const double cl = dadd(x, dsub(lo, x)); return dadd(cl, dsub(x, hi)); }
    return dadd(lo, dsub(x, hi));
  }
  if (code == 1) {
    if (D == 3) { const double cl = dadd(x, dsub(lo, x)); return dsub(cl, dsub(x, lo)); }
    return dsub(lo, dsub(x, lo));
  }
  if (D == 3) { const double cl = dadd(x, dsub(hi, x)); return dsub(cl, dsub(x, hi)); }
  return dsub(hi, dsub(x, hi));
}

template <int D, typename R = double>
__device__ __forceinline__ R hdet_iso(R h) {
  // GeomSymmetricTensor Determinant with zero off-diagonals
  if (D == 1) return h;
  if (D == 2) return h*h;
  return (h*h)*h;
}

// Largest eigenvalue of hinv * I by the reference's eigenValues formulas
// (Geometry/GeomSymmetricTensorInline.hh:2197-2256).
template <int D>
__device__ __forceinline__ double max_eigen_iso(double h) {
  if (D < 3) return h;
  const double fscale = fmax(10.0*2.220446049250313e-16, fabs(h));
  const double fi = 1.0/fscale;
  const double a = h*fi;
  const double onethird = 1.0/3.0;
  const double c0 = a*a*a;
  const double c1 = a*a + a*a + a*a;
  const double c2 = a + a + a;
  const double c2Div3 = c2*onethird;
  const double aDiv3 = fmin(0.0, onethird*(c1 - c2*c2Div3));
  const double mbDiv2 = 0.5*(c0 + c2Div3*(2.0*c2Div3*c2Div3 - c1));
  const double q = fmin(0.0, mbDiv2*mbDiv2 + aDiv3*aDiv3*aDiv3);
  const double mag = sqrt(-aDiv3);
  const double ang = atan2(sqrt(-q), mbDiv2)*onethird;
  const double cs = cos(ang), sn = sin(ang);
  const double s3 = 1.7320508075688772;
  const double e0 = fscale*(c2Div3 + 2.0*mag*cs);
  const double e1 = fscale*(c2Div3 - mag*(cs + s3*sn));
  const double e2 = fscale*(c2Div3 - mag*(cs - s3*sn));
  return fmax(e0, fmax(e1, e2));
}

// Reference rdiv: Tensor/Vector /= s multiplies by 1/s in 2-D and 3-D, divides in 1-D.
template <int D>
__device__ __forceinline__ double rdiv(double x, double s) { return D == 1 ? x/s : x*(1.0/s); }

template <typename R> __device__ __forceinline__ const R* table_coef(const DevTable& t);
template <> __device__ __forceinline__ const double* table_coef<double>(const DevTable& t) { return t.coef; }
template <> __device__ __forceinline__ const float* table_coef<float>(const DevTable& t) { return t.coef32; }

// TableKernelView::kernelAndGradValue (Kernel/TableKernelViewInline.hh:84-99) in arithmetic R.
template <typename R = double>
__device__ __forceinline__ void kernel_wgw(const DevTable& t, R eta, R hdet, R& w, R& gw) {
  if (eta < (R)t.extent) {
    int b = (int)(eta*(R)t.inv_xstep);
    b = b < t.n1 ? b : t.n1;
    const R* c = table_coef<R>(t) + 6*b;
    w = hdet*(c[0] + (c[1] + c[2]*eta)*eta);
    gw = hdet*(c[3] + (c[4] + c[5]*eta)*eta);
  } else {
    w = (R)0;
    gw = (R)0;
  }
}

// A pointer into the block's LDS, typed as such (address space 3) so that loads through it are
// LDS reads wherever the optimiser moves them.
template <typename R>
using lds_ptr = const R __attribute__((address_space(3)))*;

// The same lookup on a given copy of the coefficients (a global pointer, or an lds_ptr to the
// block's staged copy); graw is the gradient polynomial before the Hdet factor, so |graw| is
// kernelValueSPH at the same eta.  Branch-free: the interval index is clamped to the last one (in
// bounds for any eta >= 0, the float-to-int conversion saturates), and the values beyond the
// extent are selected to zero.
template <typename R, typename CP>
__device__ __forceinline__ void kernel_wgw_c(CP coef, const DevTable& t, R eta, R hdet, R& w, R& gw, R& graw) {
  int b = (int)(eta*(R)t.inv_xstep);
  b = b < t.n1 ? b : t.n1;
  const CP c = coef + 6*b;
  const bool in = eta < (R)t.extent;
  const R wv = hdet*(c[0] + (c[1] + c[2]*eta)*eta);
  const R gr = c[3] + (c[4] + c[5]*eta)*eta;
  w = in ? wv : (R)0;
  graw = in ? gr : (R)0;
  gw = in ? hdet*gr : (R)0;
}

// The kernel table's coefficients for a pair pass, staged in the block's dynamic LDS (every
// thread of the block must call this before any returns).  Pair passes call their node body
// through with_table: once with this LDS pointer and once with the global table, never with a
// generic pointer that may be either -- a load through such a pointer is a flat load, and waiting
// for its result also waits for every global load issued before it (e.g. the next neighbour's
// prefetched record).
template <typename R>
__device__ __forceinline__ lds_ptr<R> stage_table_lds(const DevTable& tab) {
  extern __shared__ __align__(16) unsigned char spx_lds[];
  const R* coef = table_coef<R>(tab);
  R* t = reinterpret_cast<R*>(spx_lds);
  const int nc = 6*(tab.n1 + 1);
  for (int k = threadIdx.x; k < nc; k += blockDim.x) t[k] = coef[k];
  __syncthreads();
  return (lds_ptr<R>)t;
}

// body(coef) with the table staged in LDS when lds != 0 (a uniform branch), else in global
// memory; body is a generic lambda, instantiated once per pointer type.
template <typename R, typename F>
__device__ __forceinline__ void with_table(const DevTable& tab, int lds, F&& body) {
  if (lds) body(stage_table_lds<R>(tab));
  else body(table_coef<R>(tab));
}

// kernelValueSPH: |dW/deta| at Hdet = 1 (Kernel/TableKernelViewInline.hh:112-120)
template <typename R = double>
__device__ __forceinline__ R kernel_wsph(const DevTable& t, R eta) {
  if (eta < (R)t.extent) {
    int b = (int)(eta*(R)t.inv_xstep);
    b = b < t.n1 ? b : t.n1;
    const R* c = table_coef<R>(t) + 6*b;
    return fabs(c[3] + (c[4] + c[5]*eta)*eta);
  }
  return (R)0;
}

// equivalentNodesPerSmoothingScale: cubic Hermite lookup (CubicHermiteInterpolatorViewInline.hh:9-33,120-125)
__device__ __forceinline__ double nperh_lookup(const DevTable& t, double x) {
  const double* v = t.herm;
  const int N = t.hn;
  double r;
  if (x < t.hxmin) {
    r = v[0] + v[N]*(x - t.hxmin);
  } else if (x > t.hxmax) {
    r = v[N - 1] + v[2*N - 1]*(x - t.hxmin);
  } else {
    int i0 = (int)(fmax(0.0, x - t.hxmin)/t.hxstep);
    i0 = i0 < N - 2 ? i0 : N - 2;
    const double tt = fmax(0.0, fmin(1.0, (x - t.hxmin - i0*t.hxstep)/t.hxstep));
    const double t2 = tt*tt, t3 = tt*t2;
    r = ((2.0*t3 - 3.0*t2 + 1.0)*v[i0] + (-2.0*t3 + 3.0*t2)*v[i0 + 1] +
         t.hxstep*((t3 - 2.0*t2 + tt)*v[N + i0] + (t3 - t2)*v[N + i0 + 1]));
  }
  return fmax(0.0, r);
}

template <int D>
__device__ __forceinline__ double rootnu(double x) {
  if (D == 1) return x;
  if (D == 2) return sqrt(x);
  return (x < 0.0 ? -1.0 : 1.0)*pow(fabs(x), 0.3333333333333333);
}

__device__ __forceinline__ uint64_t spread_bits(uint64_t v, int D) {
  uint64_t r = 0;
  for (int b = 0; b < 21; ++b) r |= ((v >> b) & 1ull) << (D*b);
  return r;
}

}  // namespace spx
