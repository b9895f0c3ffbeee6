// spx_kernels.hpp -- HIP kernels of the SPH hot path (FP64 path).
//
// Every pair pass is a gather: one lane per target particle walks its ELLPACK
// neighbour list and accumulates only its own sums, so no atomics touch the
// conserved quantities.  Where a pair term must be bitwise antisymmetric (the
// pairwise acceleration and the compatible-energy partition) both lanes evaluate
// the pair in one canonical (a, b) order: an image partner makes the gathering
// node "a" (the reference's internal-ghost pair), otherwise the lower slot is "a".
#pragma once

#include "spx_device.hpp"

namespace spx {

// ---------------------------------------------------------------------------
// Neighbour image of slot j seen from slot i: position and velocity transforms
// (PeriodicBoundary copies, ReflectingBoundary::reflectOperator flips the normal
// component: Boundary/ReflectingBoundary.cc:183-197).
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void image_pos(const Geo& g, int code, double* xj) {
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const int c = (code >> (2*a)) & 3;
    if (c) xj[a] = image_coord<D>(((g.permask >> a) & 1) ? BC_PERIODIC : BC_REFLECTING, c, xj[a], g.xmin[a], g.xmax[a]);
  }
}
template <int D>
__device__ __forceinline__ double refl_sign(const Geo& g, int code, int a) {
  const int c = (code >> (2*a)) & 3;
  if (!c) return 1.0;
  return ((g.reflmask >> (2*a + c - 1)) & 1) ? -1.0 : 1.0;
}

// ---------------------------------------------------------------------------
// Boundary enforcement + cell keys.  Periodic violators are mapped through the
// opposite plane, reflecting violators mirrored with their velocity
// (Boundary/PlanarBoundary.cc:153-200, ReflectingBoundary.cc:465-473).
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(kBlock) k_keys(Fields f, Geo g, int n, int enforce, uint64_t* keys, uint32_t* vals) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t key = 0;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double xa = f.x[a][i];
    if (enforce) {
      for (int side = 1; side <= 2; ++side) {
        const int bc = side == 1 ? g.bc_lo[a] : g.bc_hi[a];
        if (bc == BC_NONE) continue;
        const double thpt = side == 1 ? (xa - g.xmin[a]) : -(xa - g.xmax[a]);
        const bool fuzzy0 = fabs(thpt) <= 1.0e-15*fmax(1.0, fabs(thpt));
        if (!fuzzy0 && thpt < 0.0) {
          if (bc == BC_PERIODIC) {
            xa = image_coord<D>(bc, side == 1 ? 2 : 1, xa, g.xmin[a], g.xmax[a]);
          } else {
            xa = image_coord<D>(bc, side, xa, g.xmin[a], g.xmax[a]);
            f.v[a][i] = -f.v[a][i];
          }
        }
      }
      f.x[a][i] = xa;
    }
    int c = (int)floor((xa - g.origin[a])*g.inv_cell[a]);
    c = c < 0 ? 0 : (c >= g.nc[a] ? g.nc[a] - 1 : c);
    key |= spread_bits((uint64_t)c, D) << a;
  }
  keys[i] = key;
  vals[i] = (uint32_t)i;
}

// Gather a set of double arrays through a permutation (dst[p] = src[perm[p]]).
struct GatherSet {
  const double* src[48];
  double* dst[48];
  int count;
};
__global__ void __launch_bounds__(kBlock) k_gather(GatherSet s, const uint32_t* perm, int n) {
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t q = perm[p];
  for (int k = 0; k < s.count; ++k) s.dst[k][p] = s.src[k][q];
}
__global__ void __launch_bounds__(kBlock) k_gather_i64(const int64_t* src, int64_t* dst, const uint32_t* perm, int n) {
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p < n) dst[p] = src[perm[p]];
}

template <int D>
__device__ __forceinline__ int lin_cell(const Geo& g, const int* c) {
  int l = c[D - 1];
  for (int a = D - 2; a >= 0; --a) l = l*g.nc[a] + c[a];
  return l;
}
template <int D>
__device__ __forceinline__ void unspread(uint64_t key, int* c) {
  for (int a = 0; a < D; ++a) {
    uint64_t v = 0;
    for (int b = 0; b < 21; ++b) v |= ((key >> (D*b + a)) & 1ull) << b;
    c[a] = (int)v;
  }
}

// Cell ranges of the sorted keys: start[cell], end[cell] (both 0 when empty).
template <int D>
__global__ void __launch_bounds__(kBlock) k_cell_bounds(const uint64_t* keys, int n, Geo g, int* start, int* end) {
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint64_t k = keys[p];
  int c[3];
  if (p == 0 || keys[p - 1] != k) { unspread<D>(k, c); start[lin_cell<D>(g, c)] = p; }
  if (p == n - 1 || keys[p + 1] != k) { unspread<D>(k, c); end[lin_cell<D>(g, c)] = p + 1; }
}

// ---------------------------------------------------------------------------
// ConnectivityMap::computeConnectivity (Neighbor/ConnectivityMap.cc:905-947):
// j is a neighbour of i iff |Hi rij|^2 <= ext^2 or |Hj rij|^2 <= ext^2, i != j
// (an image of i itself counts, as the reference's ghost of i does).  The test is
// evaluated with the reference's rounding (no contraction), so sets match exactly.
// Also records nodeScale = 1/(lambda_max(H) nPerh) for the dt vote.
// sorted: slot of the p-th particle in cell order, or nullptr when slots are sorted.
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(kBlock) k_neighbors(Fields f, Geo g, Phys ph, int n, const int* cstart,
                                                      const int* cend, const uint32_t* sorted, Scalars* sc) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  double xi[D];
  int ci[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    xi[a] = f.x[a][i];
    int c = (int)floor((xi[a] - g.origin[a])*g.inv_cell[a]);
    ci[a] = c < 0 ? 0 : (c >= g.nc[a] ? g.nc[a] - 1 : c);
  }
  const double hi = f.hinv[i];
  f.ns[i] = 1.0/max_eigen_iso<D>(hi)/ph.nperh;
  int count = 0;
  const int cap = f.cap;
  const int r1 = D >= 2 ? 1 : 0, r2 = D >= 3 ? 1 : 0;
  for (int dz = -r2; dz <= r2; ++dz)
    for (int dy = -r1; dy <= r1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int off[3] = {dx, dy, dz};
        int cc[3] = {0, 0, 0};
        int code = 0;
        bool skip = false;
#pragma unroll
        for (int a = 0; a < D; ++a) {
          int c = ci[a] + off[a];
          if (c < 0) {
            if (g.bc_lo[a] == BC_PERIODIC) { c += g.nc[a]; code |= 1 << (2*a); }
            else if (g.bc_lo[a] == BC_REFLECTING) { c = -c - 1; code |= 1 << (2*a); }
            else skip = true;
          } else if (c >= g.nc[a]) {
            if (g.bc_hi[a] == BC_PERIODIC) { c -= g.nc[a]; code |= 2 << (2*a); }
            else if (g.bc_hi[a] == BC_REFLECTING) { c = 2*g.nc[a] - 1 - c; code |= 2 << (2*a); }
            else skip = true;
          }
          cc[a] = c;
        }
        if (skip) continue;
        const int cell = lin_cell<D>(g, cc);
        const int p0 = cstart[cell], p1 = cend[cell];
        for (int p = p0; p < p1; ++p) {
          const int j = sorted ? (int)sorted[p] : p;
          if (j == i && code == 0) continue;
          double xj[D];
#pragma unroll
          for (int a = 0; a < D; ++a) xj[a] = f.x[a][j];
          image_pos<D>(g, code, xj);
          const double hj = f.hinv[j];
          double e2i = 0.0, e2j = 0.0;
#pragma unroll
          for (int a = 0; a < D; ++a) {
            const double r = dsub(xi[a], xj[a]);
            const double ei = dmul(hi, r), ej = dmul(hj, r);
            e2i = a == 0 ? dmul(ei, ei) : dadd(e2i, dmul(ei, ei));
            e2j = a == 0 ? dmul(ej, ej) : dadd(e2j, dmul(ej, ej));
          }
          if (e2i <= ph.ext2 || e2j <= ph.ext2) {
            if (count < cap) f.nbr[(size_t)count*f.ld + i] = (uint32_t)j | ((uint32_t)code << kImgShift);
            ++count;
          }
        }
      }
  f.cnt[i] = count;
  if (count > cap) atomicMax(&sc->overflow, count);
}

// ---------------------------------------------------------------------------
// computeSPHSumMassDensity (SPH/computeSPHSumMassDensity.cc:35-85) in gather form
// -- node i receives m_j W(|Hj ri - Hj rj|, det Hj) -- fused with the GenericHydro::dt
// vote (Physics/GenericHydro.cc:112-381): per-node limits on the new density and the
// pairwise velocity-difference limit, min-reduced per block.
// ---------------------------------------------------------------------------
template <int D, bool DENS, bool DT>
__global__ void __launch_bounds__(kBlock) k_density_dt(Fields f, Geo g, Phys ph, DevTable tab, int n, double* blockMin) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  double dtmin = 1.7976931348623157e308;
  if (i < n) {
    double xi[D], vi[D];
#pragma unroll
    for (int a = 0; a < D; ++a) { xi[a] = f.x[a][i]; vi[a] = f.v[a][i]; }
    const double nsi = f.ns[i];
    double rho = 0.0;
    if (DENS) {
      const double hi = f.hinv[i];
      // self term: W(|Hi ri - Hi ri|) = W(0)
      double w, gw;
      kernel_wgw(tab, 0.0, hdet_iso<D>(hi), w, gw);
      rho = f.m[i]*w;
    }
    const int nn = f.cnt[i];
    for (int k = 0; k < nn; ++k) {
      const uint32_t e = f.nbr[(size_t)k*f.ld + i];
      const int j = (int)(e & kIdxMask);
      const int code = (int)(e >> kImgShift);
      if (DENS) {
        double xj[D];
#pragma unroll
        for (int a = 0; a < D; ++a) xj[a] = f.x[a][j];
        image_pos<D>(g, code, xj);
        const double hj = f.hinv[j];
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) { const double d = hj*xi[a] - hj*xj[a]; s += d*d; }
        double w, gw;
        kernel_wgw(tab, sqrt(s), hdet_iso<D>(hj), w, gw);
        rho += f.m[j]*w;
      }
      if (DT) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) { const double d = vi[a] - refl_sign<D>(g, code, a)*f.v[a][j]; s += d*d; }
        const double vmag = sqrt(s);
        const double d = fmin(nsi, f.ns[j])*(1.0/fmax(2.220446049250313e-16, vmag));
        dtmin = fmin(dtmin, d);
      }
    }
    if (DENS) f.rho[i] = rho; else rho = f.rho[i];
    if (DT) {
      const double tiny = 2.220446049250313e-16;
      dtmin = fmin(dtmin, nsi/(f.cs[i] + tiny));
      const double csq = sqrt(f.maxQ[i]/rho);
      dtmin = fmin(dtmin, nsi/(csq + tiny));
      double tr = 0.0;
#pragma unroll
      for (int a = 0; a < D; ++a) tr += f.DvDx[(a*D + a)*(size_t)f.ld + i];
      dtmin = fmin(dtmin, 1.0/(fabs(tr) + tiny));
      double v2 = 0.0, a2 = 0.0;
#pragma unroll
      for (int a = 0; a < D; ++a) { v2 += vi[a]*vi[a]; a2 += f.DvDt[a][i]*f.DvDt[a][i]; }
      const double vmag = sqrt(v2);
      dtmin = fmin(dtmin, 0.1*fmax(nsi/(vmag + tiny), vmag/(sqrt(a2) + tiny)));
      if (ph.useVmagDt) dtmin = fmin(dtmin, nsi/(vmag + 1.0e-10));
    }
  }
  if (DT) {
    __shared__ double red[kBlock];
    red[threadIdx.x] = dtmin;
    __syncthreads();
    for (int s = kBlock/2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) blockMin[blockIdx.x] = red[0];
  }
}

// Integrator::selectDt (Integrator/Integrator.cc:116-169) on the device.
__global__ void k_select_dt(const double* blockMin, int nblocks, Scalars* sc, Phys ph, double dtMin, double dtMax,
                            double maxTime, double dtGrowth, int useGravity) {
  __shared__ double red[kBlock];
  double m = 1.7976931348623157e308;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) m = fmin(m, blockMin[b]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kBlock/2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double hydro = red[0]*ph.cfl;
    sc->hydroDt = hydro;
    const double lo = fmin(dtMin, maxTime - sc->t), hi = fmin(dtMax, maxTime - sc->t);
    double dt = hi;
    if (hydro > 0.0 && hydro < dt) dt = hydro;
    if (useGravity) {
      const double gd = sc->gravDt;
      if (gd > 0.0 && gd < dt) dt = gd;
    }
    dt = fmin(dt, dtGrowth*sc->lastDt);
    dt = fmin(hi, fmax(lo, dt));
    sc->dt = dt;
    sc->hdt = 0.5*dt;
  }
}

// ---------------------------------------------------------------------------
// CheapSynchronousRK2 mid-step: state.update with timeAdvanceOnly (increments with
// the previous step's derivatives), then the EOS policies
// (Integrator/CheapSynchronousRK2.cc:74-77; DataBase/IncrementStateInline.hh:42-80,
// IncrementBoundedStateInline.hh; ReplaceBoundedState::updateAsIncrement; GammaLawGas).
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(kBlock) k_midstep(Fields f, Phys ph, const Scalars* sc, int n) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double hdt = sc->hdt;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    f.xm[a][i] = f.x[a][i] + hdt*f.DxDt[a][i];
    double va = f.v[a][i] + hdt*f.DvDt[a][i];
    if (ph.gravity) va = va + hdt*f.DvDtG[a][i];
    f.vm[a][i] = va;
  }
  const double rho = fmin(ph.rho_max, fmax(ph.rho_min, f.rho[i] + hdt*f.DrhoDt[i]));
  const double eps = f.eps[i] + hdt*f.DepsDt[i];
  const double h = fmin(ph.hminInv, fmax(ph.hmaxInv, f.hinv[i] + hdt*f.DHDt[i]));
  f.rhom[i] = rho;
  f.epsm[i] = eps;
  f.hinvm[i] = h;
  f.Pm[i] = ph.g1*rho*eps;
  f.csm[i] = sqrt(fmax(0.0, ph.gamma*ph.g1*eps));
}

// EOS policies on the state arrays (PressurePolicy, SoundSpeedPolicy; GammaLawGas.cc:185-245).
__global__ void __launch_bounds__(kBlock) k_eos(Fields f, Phys ph, int n) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  f.P[i] = ph.g1*f.rho[i]*f.eps[i];
  f.cs[i] = sqrt(fmax(0.0, ph.gamma*ph.g1*f.eps[i]));
}

// ---------------------------------------------------------------------------
// computeSPHOmegaGradhCorrection (SPH/computeSPHOmegaGradhCorrection.cc:136-204), gather form.
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(kBlock) k_omega(Fields f, Geo g, DevTable tab, int n, double* const* xs,
                                                  const double* hs, double* omega) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  double xi[D];
#pragma unroll
  for (int a = 0; a < D; ++a) xi[a] = xs[a][i];
  const double hi = hs[i];
  const double hdeti = hdet_iso<D>(hi);
  double wsum = 0.0, gsum = 0.0;
  const int nn = f.cnt[i];
  for (int k = 0; k < nn; ++k) {
    const uint32_t e = f.nbr[(size_t)k*f.ld + i];
    const int j = (int)(e & kIdxMask);
    const int code = (int)(e >> kImgShift);
    double xj[D];
#pragma unroll
    for (int a = 0; a < D; ++a) xj[a] = xs[a][j];
    image_pos<D>(g, code, xj);
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) { const double r = hi*(xi[a] - xj[a]); s += r*r; }
    const double eta = sqrt(s);
    double w, gw;
    kernel_wgw(tab, eta, hdeti, w, gw);
    wsum += w;
    gsum += eta*gw;
  }
  double om;
  if (nn == 0) {
    om = 1.0;
  } else {
    wsum += hdeti*tab.W0;
    om = fmax(1.0e-30, -gsum/(double(D)*wsum));
  }
  omega[i] = om;
}

}  // namespace spx
