// spx_sph.hpp -- the SPH derivative pass and the compatible-energy full update (FP64 path).
#pragma once

#include "spx_kernels.hpp"

namespace spx {

// Neighbour record at the midpoint state, image-transformed.
template <int D>
struct Nbr {
  double x[D], v[D];
  double m, rho, P, cs, h, somega;
};

template <int D>
__device__ __forceinline__ void load_mid(const Fields& f, const Geo& g, int j, int code, Nbr<D>& q) {
#pragma unroll
  for (int a = 0; a < D; ++a) { q.x[a] = f.xm[a][j]; q.v[a] = refl_sign<D>(g, code, a)*f.vm[a][j]; }
  image_pos<D>(g, code, q.x);
  q.m = f.m[j];
  q.rho = f.rhom[j];
  q.P = f.Pm[j];
  q.cs = f.csm[j];
  q.h = f.hinvm[j];
  const double om = f.omegam[j];
  q.somega = om/(om*om + 1.0e-30);           // safeInv(omega, 1e-30)
}

// Per-pair quantities of SPH::evaluateDerivativesImpl for the canonical pair (a, b)
// (SPH/SPH.cc:362-430, MonaghanGingoldViscosity::QPiij MonaghanGingoldViscosity.cc:64-100).
template <int D>
struct PairEval {
  double rab[D], vab[D], etaA[D], etaB[D], gWa[D], gWb[D], dd[D];
  double Wa, Wb, Qa, Qb, workQa, workQb, PrhoA, PrhoB, etaMagA, etaMagB;
};

template <int D>
__device__ __forceinline__ void eval_pair(const DevTable& tab, const Phys& ph, const Nbr<D>& A, const Nbr<D>& B,
                                          PairEval<D>& e) {
  double e2a = 0.0, e2b = 0.0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    e.rab[c] = A.x[c] - B.x[c];
    e.etaA[c] = A.h*e.rab[c];
    e.etaB[c] = B.h*e.rab[c];
    e2a += e.etaA[c]*e.etaA[c];
    e2b += e.etaB[c]*e.etaB[c];
    e.vab[c] = A.v[c] - B.v[c];
  }
  e.etaMagA = sqrt(e2a);
  e.etaMagB = sqrt(e2b);
  const double ia = 1.0/fmax(1.0e-30, e.etaMagA), ib = 1.0/fmax(1.0e-30, e.etaMagB);   // safeInvVar
  double gwa, gwb;
  kernel_wgw(tab, e.etaMagA, hdet_iso<D>(A.h), e.Wa, gwa);
  kernel_wgw(tab, e.etaMagB, hdet_iso<D>(B.h), e.Wb, gwb);
  const double sa = gwa*A.h, sb = gwb*B.h;                  // (gW*H)*etaUnit
  double mua = 0.0, mub = 0.0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    e.gWa[c] = sa*(e.etaA[c]*ia);
    e.gWb[c] = sb*(e.etaB[c]*ib);
    mua += e.vab[c]*e.etaA[c];
    mub += e.vab[c]*e.etaB[c];
  }
  mua = mua/(e2a + 1.0e-2);
  mub = mub/(e2b + 1.0e-2);
  const double mma = fmin(0.0, mua), mmb = fmin(0.0, mub);
  const double ea = -ph.Cl*A.cs*mma + ph.Cq*(mma*mma);
  const double eb = -ph.Cl*B.cs*mmb + ph.Cq*(mmb*mmb);
  const double QPa = ea/A.rho, QPb = eb/B.rho;
  e.Qa = A.rho*ea;
  e.Qb = B.rho*eb;
  e.PrhoA = A.somega*A.P/(A.rho*A.rho);
  e.PrhoB = B.somega*B.P/(B.rho*B.rho);
  double wqa = 0.0, wqb = 0.0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const double qa = 0.5*(QPa*e.gWa[c]), qb = 0.5*(QPb*e.gWb[c]);
    wqa += e.vab[c]*qa;
    wqb += e.vab[c]*qb;
    e.dd[c] = ((e.PrhoA*e.gWa[c] + e.PrhoB*e.gWb[c]) + qa) + qb;
  }
  e.workQa = wqa;
  e.workQb = wqb;
}

// ---------------------------------------------------------------------------
// SPH::evaluateDerivatives + SPHSmoothingScale::evaluateDerivatives on the midpoint
// state (SPH/SPH.cc:271-551, SmoothingScale/SPHSmoothingScale.cc:154-245), gather form.
// FULL also produces the outputs the integration does not consume (sum density,
// normalisation, effective Q, density gradient, first moment, M).
// ---------------------------------------------------------------------------
template <int D, bool FULL>
__global__ void __launch_bounds__(kBlock) k_sph(Fields f, Geo g, Phys ph, DevTable tab, int n) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t ld = (size_t)f.ld;
  Nbr<D> S;
#pragma unroll
  for (int a = 0; a < D; ++a) { S.x[a] = f.xm[a][i]; S.v[a] = f.vm[a][i]; }
  S.m = f.m[i]; S.rho = f.rhom[i]; S.P = f.Pm[i]; S.cs = f.csm[i]; S.h = f.hinvm[i];
  { const double om = f.omegam[i]; S.somega = om/(om*om + 1.0e-30); }

  double dvdt[D], dvdx[D*D], Mc[D*D], depsdt = 0.0, maxq = 0.0, m0 = 0.0;
  double rhosum = 0.0, nrm = 0.0, effq = 0.0, grho[D], m1[D];
#pragma unroll
  for (int a = 0; a < D; ++a) { dvdt[a] = 0.0; grho[a] = 0.0; m1[a] = 0.0; }
#pragma unroll
  for (int a = 0; a < D*D; ++a) { dvdx[a] = 0.0; Mc[a] = 0.0; }

  const int nn = f.cnt[i];
  for (int k = 0; k < nn; ++k) {
    const uint32_t ent = f.nbr[(size_t)k*ld + i];
    const int j = (int)(ent & kIdxMask);
    const int code = (int)(ent >> kImgShift);
    Nbr<D> Q;
    load_mid<D>(f, g, j, code, Q);
    const bool selfA = code != 0 || i < j;
    PairEval<D> e;
    if (selfA) eval_pair<D>(tab, ph, S, Q, e); else eval_pair<D>(tab, ph, Q, S, e);
    // this node's side of the pair
    const double Ws = selfA ? e.Wa : e.Wb;
    const double Qs = selfA ? e.Qa : e.Qb;
    const double wQs = selfA ? e.workQa : e.workQb;
    const double Prs = selfA ? e.PrhoA : e.PrhoB;
    const double mo = Q.m, rhoo = Q.rho;
    const double dsg = selfA ? -1.0 : 1.0;
    double vdg = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const double gs = selfA ? e.gWa[c] : e.gWb[c];
      dvdt[c] += dsg*(mo*e.dd[c]);
      vdg += e.vab[c]*gs;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const double gd = selfA ? e.gWa[d] : e.gWb[d];
        dvdx[c*D + d] -= mo*(e.vab[c]*gd);
        Mc[c*D + d] -= mo*(e.rab[c]*gd);
      }
    }
    depsdt += mo*(Prs*vdg + wQs);
    maxq = fmax(maxq, Qs);
    const double wsph = kernel_wsph(tab, selfA ? e.etaMagA : e.etaMagB);
    m0 += wsph;
    if (FULL) {
      rhosum += mo*Ws;
      nrm += S.m/S.rho*Ws;
      effq += mo*Qs*Ws/rhoo;
      const double drho = selfA ? (rhoo - S.rho) : (S.rho - rhoo);
#pragma unroll
      for (int c = 0; c < D; ++c) {
        grho[c] += mo*drho*(selfA ? e.gWa[c] : e.gWb[c]);
        m1[c] += dsg*(wsph*(selfA ? e.etaA[c] : e.etaB[c]));
      }
    }
  }

  // ---- per-node finalize (SPH.cc:480-551)
  const double hdet = hdet_iso<D>(S.h);
  double det;
  if (D == 1) det = Mc[0];
  else if (D == 2) det = Mc[0]*Mc[3] - Mc[1]*Mc[2];
  else det = (Mc[0]*Mc[4]*Mc[8] + Mc[1]*Mc[5]*Mc[6] + Mc[2]*Mc[3]*Mc[7] -
              Mc[0]*Mc[5]*Mc[7] - Mc[1]*Mc[3]*Mc[8] - Mc[2]*Mc[4]*Mc[6]);
  const double pow2d = D == 1 ? 2.0 : (D == 2 ? 4.0 : 8.0);
  double G[D*D];
  if (ph.correctVG && fabs(det) > 1.0e-10 && (double)nn > pow2d) {
    double Mi[D*D];
    if (D == 1) {
      Mi[0] = 1.0/Mc[0];
    } else if (D == 2) {
      const double r = 1.0/det;
      Mi[0] = Mc[3]*r; Mi[1] = -Mc[1]*r; Mi[2] = -Mc[2]*r; Mi[3] = Mc[0]*r;
    } else {
      const double r = 1.0/det;
      const double xx = Mc[0], xy = Mc[1], xz = Mc[2], yx = Mc[3], yy = Mc[4], yz = Mc[5], zx = Mc[6], zy = Mc[7], zz = Mc[8];
      Mi[0] = (yy*zz - yz*zy)*r; Mi[1] = (xz*zy - xy*zz)*r; Mi[2] = (xy*yz - xz*yy)*r;
      Mi[3] = (yz*zx - yx*zz)*r; Mi[4] = (xx*zz - xz*zx)*r; Mi[5] = (xz*yx - xx*yz)*r;
      Mi[6] = (yx*zy - yy*zx)*r; Mi[7] = (xy*zx - xx*zy)*r; Mi[8] = (xx*yy - xy*yx)*r;
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) {
        double s = 0.0;
#pragma unroll
        for (int k2 = 0; k2 < D; ++k2) s += dvdx[a*D + k2]*Mi[k2*D + b];
        G[a*D + b] = s;
      }
    if (FULL) {
#pragma unroll
      for (int a = 0; a < D*D; ++a) Mc[a] = Mi[a];
    }
  } else {
#pragma unroll
    for (int a = 0; a < D*D; ++a) G[a] = rdiv<D>(dvdx[a], S.rho);
  }
  double tr = 0.0;
#pragma unroll
  for (int a = 0; a < D; ++a) tr += G[a*D + a];
#pragma unroll
  for (int a = 0; a < D*D; ++a) f.DvDx[a*ld + i] = G[a];
#pragma unroll
  for (int a = 0; a < D; ++a) { f.DvDt[a][i] = dvdt[a]; f.DxDt[a][i] = S.v[a]; }
  f.DepsDt[i] = depsdt;
  f.DrhoDt[i] = -S.rho*tr;
  f.maxQ[i] = maxq;
  // ---- SPHSmoothingScale finalize (SPHSmoothingScale.cc:228-245)
  m0 = rootnu<D>(fmax(0.0, m0));
  f.m0[i] = m0;
  f.DHDt[i] = rdiv<D>(-S.h, double(D))*tr;
  const double cur = (fabs(m0) <= 1.0e-15*fmax(1.0, fabs(m0))) ? 0.5*ph.nperh : nperh_lookup(tab, m0);
  const double s = fmin(4.0, fmax(0.25, ph.nperh/(cur + 1.0e-30)));
  const double aa = s < 1.0 ? 0.4*(1.0 + s*s) : 0.4*(1.0 + 1.0/(s*s*s));
  const double hi0 = 1.0/S.h;
  const double hi1 = fmin(ph.hmax, fmax(ph.hmin, hi0*(1.0 - aa + aa*s)));
  f.Hideal[i] = 1.0/hi1;
  if (FULL) {
    f.rhoSum[i] = rhosum + S.m*tab.W0*hdet;
    f.norm[i] = nrm + S.m/S.rho*tab.W0*hdet;
    f.effQ[i] = effq;
#pragma unroll
    for (int c = 0; c < D; ++c) { f.gradRho[c][i] = rdiv<D>(grho[c], S.rho); f.m1[c][i] = m1[c]; }
#pragma unroll
    for (int a = 0; a < D*D; ++a) f.Mcorr[a*ld + i] = Mc[a];
  }
}

// ---------------------------------------------------------------------------
// Full step from the start-of-step state with the midpoint derivatives
// (CheapSynchronousRK2.cc:106-109): SpecificThermalEnergyPolicy::update
// (Hydro/SpecificThermalEnergyPolicy.cc:79-144) with the pair accelerations
// recomputed from the midpoint state in canonical order, then the increment and
// replace policies and the EOS.  v0 is read from f.v, the new velocity goes to
// f.v1.  ENERGY_ONLY: the package-level policy call (eps only).
// ---------------------------------------------------------------------------
template <int D, bool ENERGY_ONLY>
__global__ void __launch_bounds__(kBlock) k_full_update(Fields f, Geo g, Phys ph, DevTable tab, const Scalars* sc,
                                                        double multiplier, int n) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dt = ENERGY_ONLY ? multiplier : sc->dt;
  const double hdt = 0.5*dt;
  const size_t ld = (size_t)f.ld;
  Nbr<D> S;
#pragma unroll
  for (int a = 0; a < D; ++a) { S.x[a] = f.xm[a][i]; S.v[a] = f.vm[a][i]; }
  S.m = f.m[i]; S.rho = f.rhom[i]; S.P = f.Pm[i]; S.cs = f.csm[i]; S.h = f.hinvm[i];
  { const double om = f.omegam[i]; S.somega = om/(om*om + 1.0e-30); }
  double v12s[D];
#pragma unroll
  for (int a = 0; a < D; ++a) v12s[a] = f.v[a][i] + f.DvDt[a][i]*hdt;
  const double de0s = f.DepsDt[i];
  const double epsm = 2.220446049250313e-16;
  double du = 0.0;
  const int nn = f.cnt[i];
  for (int k = 0; k < nn; ++k) {
    const uint32_t ent = f.nbr[(size_t)k*ld + i];
    const int j = (int)(ent & kIdxMask);
    const int code = (int)(ent >> kImgShift);
    Nbr<D> Q;
    load_mid<D>(f, g, j, code, Q);
    double v12o[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const double sg = refl_sign<D>(g, code, a);
      v12o[a] = sg*f.v[a][j] + (sg*f.DvDt[a][j])*hdt;
    }
    const double de0o = f.DepsDt[j];
    const bool selfA = code != 0 || i < j;
    PairEval<D> e;
    if (selfA) eval_pair<D>(tab, ph, S, Q, e); else eval_pair<D>(tab, ph, Q, S, e);
    // pacc_ab = -m_b * deltaDvDt ; du_ab = (v_b^{1/2} - v_a^{1/2}) . pacc_ab
    const double mb = selfA ? Q.m : S.m, ma = selfA ? S.m : Q.m;
    double duab = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const double va = selfA ? v12s[c] : v12o[c], vb = selfA ? v12o[c] : v12s[c];
      duab += (vb - va)*(-mb*e.dd[c]);
    }
    const double sgn = duab < 0.0 ? -1.0 : 1.0;
    const double wa = fmax(epsm, (selfA ? de0s : de0o)*sgn), wb = fmax(epsm, (selfA ? de0o : de0s)*sgn);
    const double w = wa/(wa + wb);
    du += selfA ? w*duab : (1.0 - w)*duab*ma/mb;
  }
  const double eps1 = f.eps[i] + du*dt;
  f.eps[i] = eps1;
  if (ENERGY_ONLY) return;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    f.x[a][i] = f.x[a][i] + dt*f.DxDt[a][i];
    double va = f.v[a][i] + dt*f.DvDt[a][i];
    if (ph.gravity) va = va + dt*f.DvDtG[a][i];
    f.v1[a][i] = va;
  }
  const double rho1 = fmin(ph.rho_max, fmax(ph.rho_min, f.rho[i] + dt*f.DrhoDt[i]));
  f.rho[i] = rho1;
  f.hinv[i] = fmin(ph.hminInv, fmax(ph.hmaxInv, f.Hideal[i]));
  f.P[i] = ph.g1*rho1*eps1;
  f.cs[i] = sqrt(fmax(0.0, ph.gamma*ph.g1*eps1));
}

// Reductions for the next step's grid: min hinv and the bounding box (per block).
template <int D>
__global__ void __launch_bounds__(kBlock) k_extent(Fields f, int n, double* part) {
  __shared__ double red[kBlock];
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  double vals[1 + 2*3];
  vals[0] = i < n ? f.hinv[i] : 1.7976931348623157e308;
  for (int a = 0; a < D; ++a) {
    const double x = i < n ? f.x[a][i] : 0.0;
    vals[1 + 2*a] = i < n ? x : 1.7976931348623157e308;
    vals[2 + 2*a] = i < n ? -x : 1.7976931348623157e308;
  }
  for (int q = 0; q < 1 + 2*D; ++q) {
    red[threadIdx.x] = vals[q];
    __syncthreads();
    for (int s = kBlock/2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x*(1 + 2*D) + q] = red[0];
    __syncthreads();
  }
}

// Non-finite guard on the state (NaN/Inf -> status word).
template <int D>
__global__ void __launch_bounds__(kBlock) k_finite(Fields f, int n, Scalars* sc) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool ok = isfinite(f.rho[i]) && isfinite(f.eps[i]) && isfinite(f.hinv[i]);
  for (int a = 0; a < D; ++a) ok = ok && isfinite(f.x[a][i]) && isfinite(f.v[a][i]);
  if (!ok) atomicOr(&sc->pad, 1);
}

}  // namespace spx
