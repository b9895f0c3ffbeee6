// spx_sph.hpp -- the SPH derivative pass and the compatible-energy full update (FP64 path).
#pragma once

#include "spx_kernels.hpp"

namespace spx {

// Neighbour record at the midpoint state, image-transformed.
template <int D, typename R>
struct Nbr {
  double x[D], v[D];     // positions and velocities stay double: differences are formed before rounding to R
  R m, rho, cs, h, prho, irho;
};

// A neighbour's midpoint record as loaded (no image transform yet), so the loads of the
// next neighbour can be issued before the current pair is evaluated.
template <int D>
struct NbrRaw {
  double x[D], v[D], m, rho, cs, h, prho, irho;
};
template <int D>
__device__ __forceinline__ void load_raw(const Fields& f, int j, NbrRaw<D>& q) {
  const double2* X = reinterpret_cast<const double2*>(f.xhm) + 2*(size_t)j;
  const double2* P = reinterpret_cast<const double2*>(f.pkm) + 4*(size_t)j;
  const double2 a = X[0], b = X[1], c = P[0], d = P[1], e = P[2], p = P[3];
  q.x[0] = a.x;
  if (D >= 2) q.x[1 % D] = a.y;
  if (D >= 3) q.x[2 % D] = b.x;
  q.h = b.y;
  q.v[0] = c.x;
  if (D >= 2) q.v[1 % D] = c.y;
  if (D >= 3) q.v[2 % D] = d.x;
  q.m = d.y;
  q.rho = e.x;
  q.cs = e.y;
  q.prho = p.x;
  q.irho = p.y;
}
// Image-transformed neighbour record in the pair arithmetic type.
template <int D, typename R>
__device__ __forceinline__ void finish_mid(const Geo& g, int code, const NbrRaw<D>& r, Nbr<D, R>& q) {
#pragma unroll
  for (int a = 0; a < D; ++a) { q.x[a] = r.x[a]; q.v[a] = refl_sign<D>(g, code, a)*r.v[a]; }
  if (code) image_pos<D>(g, code, q.x);
  q.m = (R)r.m;
  q.rho = (R)r.rho;
  q.cs = (R)r.cs;
  q.h = (R)r.h;
  q.prho = (R)r.prho;
  q.irho = (R)r.irho;
}
template <int D, typename R>
__device__ __forceinline__ void load_self(const Fields& f, int i, Nbr<D, R>& s) {
#pragma unroll
  for (int a = 0; a < D; ++a) { s.x[a] = f.xm[a][i]; s.v[a] = f.vm[a][i]; }
  s.m = (R)f.m[i];
  s.rho = (R)f.rhom[i];
  s.cs = (R)f.csm[i];
  s.h = (R)f.hinvm[i];
  s.prho = (R)f.prhom[i];
  s.irho = (R)reinterpret_cast<const double2*>(f.pkm)[4*(size_t)i + 3].y;   // the value partners read
}

// Pair quantities of SPH::evaluateDerivativesImpl (SPH/SPH.cc:362-430) and
// MonaghanGingoldViscosity::QPiij (MonaghanGingoldViscosity.cc:64-100), evaluated
// from the gathering node s toward its neighbour o (r = r_s - r_o).  The canonical
// pair (a, b) of the reference has a = s when aFirst, else a = o.
//
// With isotropic H (H = h^-1 I) each side's kernel gradient H etaUnit dW is parallel to r:
// gradW_s = ks r with ks = (dW_s h_s^-1)(h_s^-1 / |eta_s|), and the Q's tensor is a scalar times
// the identity, so the pairwise acceleration (SPH.cc:428) is one scalar times r:
//   pacc_ab / (-m_b) = C r_ab,   C = ((P_a k_a + P_b k_b) + Q_a k_a / 2) + Q_b k_b / 2
// (P = safeInv(Omega) P / rho^2, Q = the Q's QPi).  C is formed in that canonical order from
// quantities both lanes compute with the same operations on the same operands (r and the
// velocity jump change sign exactly), so both lanes hold the same C bits and each lane's
// C*r_lane is the exact negation of the other's: the pairwise acceleration is bitwise
// antisymmetric.  C per neighbour entry is all the compatible-energy update needs besides the
// positions, which it re-forms with the same operations.
template <int D, typename R>
struct PairS {
  R r[D], vso[D], gs[D], es[D];
  R C, Ws, Qs, workQs, vdg, etaMagS, wsphS;
};

// coef: the kernel table's coefficients in arithmetic R (global, or the block's LDS copy).
// LQ: LimitedMonaghanGingoldViscosity::QPiij (LimitedMonaghanGingoldViscosity.cc:160-195) in
// place of the plain Q; TS, TO are the two nodes' velocity gradients as the Q holds them
// (row-major D x D, the partner's already image-transformed).
template <int D, typename R, bool LQ = false, typename CP>
__device__ __forceinline__ void pair_eval(const DevTable& tab, CP coef, const Phys& ph, const Nbr<D, R>& S,
                                          const Nbr<D, R>& O, bool aFirst, PairS<D, R>& e,
                                          const R* TS = nullptr, const R* TO = nullptr) {
  R e2s = (R)0, e2o = (R)0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    e.r[c] = (R)(S.x[c] - O.x[c]);
    e.es[c] = S.h*e.r[c];
    const R eoc = O.h*e.r[c];
    e2s += e.es[c]*e.es[c];
    e2o += eoc*eoc;
    e.vso[c] = (R)(S.v[c] - O.v[c]);
  }
  // |eta| and safeInvVar(|eta|) = 1/max(1e-30, |eta|) from one reciprocal square root each
  const R tiny = sizeof(R) == sizeof(double) ? (R)1.0e-30 : (R)1.0e-30f;
  const R is = e2s > tiny*tiny ? rsqrt(e2s) : (R)1/tiny, io = e2o > tiny*tiny ? rsqrt(e2o) : (R)1/tiny;
  e.etaMagS = e2s*is;
  const R etaMagO = e2o*io;
  R gws, gwo, graws, grawo, Wo;
  kernel_wgw_c<R>(coef, tab, e.etaMagS, hdet_iso<D, R>(S.h), e.Ws, gws, graws);
  kernel_wgw_c<R>(coef, tab, etaMagO, hdet_iso<D, R>(O.h), Wo, gwo, grawo);
  e.wsphS = fabs(graws);                                      // kernelValueSPH(etaMagS)
  const R ks = (gws*S.h)*(S.h*is), ko = (gwo*O.h)*(O.h*io);  // gradW = k r
  R vsr = (R)0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    e.gs[c] = ks*e.r[c];
    vsr += e.vso[c]*e.r[c];
  }
  // velocity jump the Q sees: the plain difference, or the limited midpoint extrapolation
  R vqr = vsr;
  if (LQ) {
    // van Leer limiter on the ratio of the gradients' projections along the half separation,
    // phased back toward the full Q when the pair is closer than etaCrit
    R xij[D], gi = (R)0, gj = (R)0;
#pragma unroll
    for (int c = 0; c < D; ++c) xij[c] = (R)0.5*e.r[c];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      R s1 = TS[a*D]*xij[0], s2 = TO[a*D]*xij[0];
#pragma unroll
      for (int b = 1; b < D; ++b) { s1 += TS[a*D + b]*xij[b]; s2 += TO[a*D + b]*xij[b]; }
      gi = a == 0 ? s1*xij[0] : gi + s1*xij[a];
      gj = a == 0 ? s2*xij[0] : gj + s2*xij[a];
    }
    const R tiny30 = sizeof(R) == sizeof(double) ? (R)1.0e-30 : (R)1.0e-30f;
    const R ri = gi/((gj < (R)0 ? (R)-1 : (R)1)*fmax(tiny30, fabs(gj)));
    const R rj = gj/((gi < (R)0 ? (R)-1 : (R)1)*fmax(tiny30, fabs(gi)));
    const R rmin = fmin(ri, rj);
    R phi = rmin > (R)0 ? (R)2/((R)1 + rmin)*(R)2*rmin/((R)1 + rmin) : (R)0;
    const R etaij = fmin(e.etaMagS, etaMagO), ec = (R)ph.etaCrit;
    if (etaij < ec) { const R q = (etaij - ec)/(R)ph.etaFold; phi *= exp(-(q*q)); }
    // (phi*T).xij: the product with the scaled tensor, as the reference forms it
    vqr = (R)0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      R s1 = (phi*TS[a*D])*xij[0], s2 = (phi*TO[a*D])*xij[0];
#pragma unroll
      for (int b = 1; b < D; ++b) { s1 += (phi*TS[a*D + b])*xij[b]; s2 += (phi*TO[a*D + b])*xij[b]; }
      const R vq = (R)((S.v[a] - (double)s1) - (O.v[a] + (double)s2));
      vqr += vq*e.r[a];
    }
  }
  // mu = (vq . eta)/(eta^2 + eps^2) with eta = h^-1 r on each side
  const R mus = (S.h*vqr)/(e2s + (R)1.0e-2), muo = (O.h*vqr)/(e2o + (R)1.0e-2);
  const R mms = fmin((R)0, mus), mmo = fmin((R)0, muo);
  const R Cl = (R)ph.Cl, Cq = (R)ph.Cq;
  const R es_ = -Cl*S.cs*mms + Cq*(mms*mms);
  const R eo_ = -Cl*O.cs*mmo + Cq*(mmo*mmo);
  const R QPs = es_*S.irho, QPo = eo_*O.irho;
  e.Qs = S.rho*es_;
  const R p1 = aFirst ? S.prho : O.prho, p2 = aFirst ? O.prho : S.prho;
  const R k1 = aFirst ? ks : ko, k2 = aFirst ? ko : ks;
  const R q1 = aFirst ? QPs : QPo, q2 = aFirst ? QPo : QPs;
  e.C = ((p1*k1 + p2*k2) + (R)0.5*(q1*k1)) + (R)0.5*(q2*k2);
  e.vdg = ks*vsr;                                            // v_so . gradW_s
  e.workQs = (R)0.5*(QPs*e.vdg);                             // v_so . (QPi_s gradW_s / 2)
}

// ---------------------------------------------------------------------------
// SPH::evaluateDerivatives + SPHSmoothingScale::evaluateDerivatives on the midpoint
// state (SPH/SPH.cc:271-551, SmoothingScale/SPHSmoothingScale.cc:154-245), gather form.
// FULL also produces the outputs the integration does not consume (sum density,
// normalisation, effective Q, density gradient, first moment, M).
// ---------------------------------------------------------------------------
// ldsTable != 0: the table's coefficients are staged in the block's dynamic LDS first (the
// pair loop makes three lookups per pair at scattered intervals).  sc != nullptr (the step):
// the node's half-step velocity record for the compatible-energy update is written too (as
// k_v12 would, with the run's dt/2).
template <typename R>
constexpr bool kSymM = sizeof(R) == sizeof(double);

// Lanes per node of the SPH pass.  With two, the lanes of an adjacent (even, odd) pair take the
// node's ELL entries k = 0, 2, 4, ... and 1, 3, 5, ... and add their partial sums at the end (a
// DPP swap within the pair: both lanes form the same a + b), doubling the pass's parallelism
// for the same registers per lane.  Measured on one MI355X (SPH pass time): KH 2-D 4096 mixed
// 57.8 -> 32.6 us, KH 2-D 65536 mixed 75.0 -> 52.7 us, TG 3-D 16^3 FP64 65.6 -> 39.8 us,
// TG 3-D 128^3 mixed 1.244 -> 1.209 ms, but TG 3-D 128^3 FP64 1.362 -> 1.384 ms and TG 3-D 48^3
// FP64 unchanged.  The choice depends only on the configuration (dimension, precision), never on
// the particle count, so a run gives the same bits in any batch.
__host__ __device__ constexpr int sph_lanes(int D, bool mixed) { return (mixed || D <= 2) ? 2 : 1; }
template <int D, typename R>
constexpr int kSphLanes = sph_lanes(D, sizeof(R) == sizeof(float));

__device__ __forceinline__ double lane_pair_swap(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  // quad_perm [1, 0, 3, 2]: lane 2m reads lane 2m+1 and vice versa
  const int lo2 = __builtin_amdgcn_mov_dpp(lo, 0xB1, 0xf, 0xf, false);
  const int hi2 = __builtin_amdgcn_mov_dpp(hi, 0xB1, 0xf, 0xf, false);
  return __hiloint2double(hi2, lo2);
}

template <int D, bool FULL, typename R, bool LQ, typename CP>
__device__ __forceinline__ void sph_node(const Fields& f, const Geo& g, const Phys& ph, const DevTable& tab,
                                         CP coef, int n, const Scalars* sc) {
  constexpr int L = kSphLanes<D, R>;
  const int gl = xcd_block()*blockDim.x + threadIdx.x;
  const int i = L == 1 ? gl : gl >> 1, sub = L == 1 ? 0 : (gl & 1);
  // both lanes of a pair share i: they leave together
  if (i >= n) return;
  const size_t ld = (size_t)f.ld;
  R* pdd = (R*)f.pdd;
  Nbr<D, R> S;
  load_self<D, R>(f, i, S);
  const double Sm = f.m[i], Srho = f.rhom[i], Sh = f.hinvm[i];

  double dvdt[D], dvdx[D*D], Mc[D*D], depsdt = 0.0, maxq = 0.0, m0 = 0.0;
  double rhosum = 0.0, nrm = 0.0, effq = 0.0, grho[D], m1[D];
#pragma unroll
  for (int a = 0; a < D; ++a) { dvdt[a] = 0.0; grho[a] = 0.0; m1[a] = 0.0; }
#pragma unroll
  for (int a = 0; a < D*D; ++a) { dvdx[a] = 0.0; Mc[a] = 0.0; }

  R TS[LQ ? D*D : 1];
  if (LQ) {
#pragma unroll
    for (int a = 0; a < D*D; ++a) TS[a] = (R)f.DvDxQ[a*ld + i];
  }
  const int nn = f.cnt[i];
  // the next ELL entry and its neighbour's record are fetched one pair ahead (one lane: the
  // entries come 8 at a time in the blocked list's 32-byte chunks)
  uint4 c0 = make_uint4(0u, 0u, 0u, 0u), c1 = c0;
  uint32_t en = 0;
  NbrRaw<D> rn;
  if (sub < nn) {
    if (L == 1) {
      const uint4* cp = reinterpret_cast<const uint4*>(f.nbr + ell_at(0, i, ld));
      c0 = cp[0];
      c1 = cp[1];
      en = c0.x;
    } else {
      en = f.nbr[ell_at(sub, i, ld)];
    }
    load_raw<D>(f, (int)(en & kIdxMask), rn);
  }
  for (int k = sub; k < nn; k += L) {
    const uint32_t ent = en;
    const NbrRaw<D> rc = rn;
    if (k + L < nn) {
      const int k1 = k + L;
      if (L == 1) {
        const int u = k1 & 7;
        if (u == 0) {
          const uint4* cp = reinterpret_cast<const uint4*>(f.nbr + ell_at(k1, i, ld));
          c0 = cp[0];
          c1 = cp[1];
        }
        en = u == 0 ? c0.x : u == 1 ? c0.y : u == 2 ? c0.z : u == 3 ? c0.w
           : u == 4 ? c1.x : u == 5 ? c1.y : u == 6 ? c1.z : c1.w;
      } else {
        en = f.nbr[ell_at(k1, i, ld)];
      }
      load_raw<D>(f, (int)(en & kIdxMask), rn);
    }
    const int j = (int)(ent & kIdxMask);
    const int code = (int)(ent >> kImgShift);
    Nbr<D, R> O;
    finish_mid<D, R>(g, code, rc, O);
    PairS<D, R> e;
    if (LQ) {
      // the partner's gradient: ghosts carry the tensor image (reflected axes flip sign per
      // index) when boundaries were applied after the Q's copy; otherwise the zero the cycle's
      // ghost rebuild left (ArtificialViscosityHandle.cc:168-183, Integrator.cc:254-271)
      R TO[D*D];
      const bool img = code != 0;
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
          const double t = f.DvDxQ[(size_t)(a*D + b)*ld + j];
          TO[a*D + b] = img ? (ph.qImages ? (R)(refl_sign<D>(g, code, a)*refl_sign<D>(g, code, b)*t) : (R)0) : (R)t;
        }
      pair_eval<D, R, true>(tab, coef, ph, S, O, code != 0 || i < j, e, TS, TO);
    } else {
      pair_eval<D, R>(tab, coef, ph, S, O, code != 0 || i < j, e);
    }
    pdd[(size_t)k*ld + i] = e.C;
    const double mo = rc.m;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      dvdt[c] -= mo*(double)(e.C*e.r[c]);
      const double mv = mo*(double)e.vso[c], mr = mo*(double)e.r[c];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        dvdx[c*D + d] -= mv*(double)e.gs[d];
        // with isotropic H the gradient is parallel to r, so M = sum m r (x) gradW is symmetric:
        // in FP64 the upper triangle is accumulated and mirrored after the loop (the FP32 pair
        // terms round r_c g_d and r_d g_c differently, so mixed precision keeps all nine)
        if (!kSymM<R> || d >= c) Mc[c*D + d] -= mr*(double)e.gs[d];
      }
    }
    depsdt += mo*(double)(S.prho*e.vdg + e.workQs);
    maxq = fmax(maxq, (double)e.Qs);
    const R wsph = e.wsphS;
    m0 += (double)wsph;
    if (FULL) {
      const double orho = rc.rho;
      rhosum += mo*(double)e.Ws;
      nrm += Sm/Srho*(double)e.Ws;
      effq += mo*(double)e.Qs*(double)e.Ws/orho;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        grho[c] += mo*(orho - Srho)*(double)e.gs[c];
        m1[c] -= (double)(wsph*e.es[c]);
      }
    }
  }

  if (L == 2) {
    // the pair's partial sums: each lane forms (own + partner's), the same bits on both
#pragma unroll
    for (int a = 0; a < D; ++a) dvdt[a] += lane_pair_swap(dvdt[a]);
#pragma unroll
    for (int c = 0; c < D; ++c)
#pragma unroll
      for (int d = 0; d < D; ++d) {
        dvdx[c*D + d] += lane_pair_swap(dvdx[c*D + d]);
        if (!kSymM<R> || d >= c) Mc[c*D + d] += lane_pair_swap(Mc[c*D + d]);
      }
    depsdt += lane_pair_swap(depsdt);
    m0 += lane_pair_swap(m0);
    maxq = fmax(maxq, lane_pair_swap(maxq));
    if (FULL) {
      rhosum += lane_pair_swap(rhosum);
      nrm += lane_pair_swap(nrm);
      effq += lane_pair_swap(effq);
#pragma unroll
      for (int a = 0; a < D; ++a) { grho[a] += lane_pair_swap(grho[a]); m1[a] += lane_pair_swap(m1[a]); }
    }
    if (sub != 0) return;
  }
  if (kSymM<R>) {
#pragma unroll
    for (int c = 1; c < D; ++c)
#pragma unroll
      for (int d = 0; d < c; ++d) Mc[c*D + d] = Mc[d*D + c];
  }
  // ---- per-node finalize (SPH.cc:480-551), double precision
  const double hdet = hdet_iso<D>(Sh);
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
    for (int a = 0; a < D*D; ++a) G[a] = rdiv<D>(dvdx[a], Srho);
  }
  double tr = 0.0;
#pragma unroll
  for (int a = 0; a < D; ++a) tr += G[a*D + a];
#pragma unroll
  for (int a = 0; a < D*D; ++a) f.DvDx[a*ld + i] = G[a];
#pragma unroll
  for (int a = 0; a < D; ++a) { f.DvDt[a][i] = dvdt[a]; f.DxDt[a][i] = f.vm[a][i]; }
  f.DepsDt[i] = depsdt;
  {
    // the partition record: this evaluation's positions (the pair coefficients' r) always; the
    // half-step velocities and DepsDt in the step (otherwise k_v12 writes them)
    double2* PV = reinterpret_cast<double2*>(f.pkv) + 4*(size_t)i;
    double xs[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < D; ++a) xs[a] = S.x[a];
    if (sc) {
      const double hdt = 0.5*sc[f.run[i]].dt;
      double w[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int a = 0; a < D; ++a) w[a] = f.v[a][i] + dvdt[a]*hdt;
      PV[0] = make_double2(w[0], w[1]);
      PV[1] = make_double2(w[2], depsdt);
    }
    PV[2] = make_double2(Sm, xs[0]);
    PV[3] = make_double2(xs[1], xs[2]);
  }
  f.DrhoDt[i] = -Srho*tr;
  f.maxQ[i] = maxq;
  // ---- SPHSmoothingScale finalize (SPHSmoothingScale.cc:228-245)
  m0 = rootnu<D>(fmax(0.0, m0));
  f.m0[i] = m0;
  f.DHDt[i] = rdiv<D>(-Sh, double(D))*tr;
  const double cur = (fabs(m0) <= 1.0e-15*fmax(1.0, fabs(m0))) ? 0.5*ph.nperh : nperh_lookup(tab, m0);
  const double s = fmin(4.0, fmax(0.25, ph.nperh/(cur + 1.0e-30)));
  const double aa = s < 1.0 ? 0.4*(1.0 + s*s) : 0.4*(1.0 + 1.0/(s*s*s));
  const double hi0 = 1.0/Sh;
  const double hi1 = fmin(ph.hmax, fmax(ph.hmin, hi0*(1.0 - aa + aa*s)));
  f.Hideal[i] = 1.0/hi1;
  if (FULL) {
    f.rhoSum[i] = rhosum + Sm*tab.W0*hdet;
    f.norm[i] = nrm + Sm/Srho*tab.W0*hdet;
    f.effQ[i] = effq;
#pragma unroll
    for (int c = 0; c < D; ++c) { f.gradRho[c][i] = rdiv<D>(grho[c], Srho); f.m1[c][i] = m1[c]; }
#pragma unroll
    for (int a = 0; a < D*D; ++a) f.Mcorr[a*ld + i] = Mc[a];
  }
}

template <int D, bool FULL, typename R, bool LQ = false>
__global__ void __launch_bounds__(kBlock, 3) k_sph(Fields f, Geo g, Phys ph, DevTable tab, int n, int ldsTable,
                                                   const Scalars* sc) {
  with_table<R>(tab, ldsTable, [&](auto coef) { sph_node<D, FULL, R, LQ>(f, g, ph, tab, coef, n, sc); });
}

// ---------------------------------------------------------------------------
// SpecificThermalEnergyPolicy::update's partition of the pair work (Hydro/
// SpecificThermalEnergyPolicy.cc:79-144): node i's specific thermal energy rate from the pair
// coefficients C the SPH pass stored per ELL entry and the partition records (Fields::pkv: the
// half-step velocity, DepsDt, m and the evaluation's position).  The pairwise acceleration is
// re-formed as C*r with r computed from the evaluation's positions by the SPH pass's own
// operations, so it is the same bits the SPH pass summed into DvDt.
// ---------------------------------------------------------------------------
template <int D, typename R, int G>
__device__ __forceinline__ double compatible_rate(const Fields& f, const Geo& g, int i) {
  const size_t ld = (size_t)f.ld;
  const R* pdd = (const R*)f.pdd;       // pair coefficients in the pair arithmetic; the partition is FP64
  const double ms = f.m[i];
  double v12s[D], xs[D];
  const double2* PV = reinterpret_cast<const double2*>(f.pkv);
  {
    const double2 a = PV[4*(size_t)i], b = PV[4*(size_t)i + 1], c = PV[4*(size_t)i + 2], d = PV[4*(size_t)i + 3];
    v12s[0] = a.x;
    if (D >= 2) v12s[1 % D] = a.y;
    if (D >= 3) v12s[2 % D] = b.x;
    xs[0] = c.y;
    if (D >= 2) xs[1 % D] = d.x;
    if (D >= 3) xs[2 % D] = d.y;
  }
  const double de0s = f.DepsDt[i];
  const double epsm = 2.220446049250313e-16;
  double du = 0.0;
  const int nn = f.cnt[i];
  for (int k0 = 0; k0 < nn; k0 += G) {
    int js[G], codes[G];
    ell_group<G>(f.nbr, ld, k0, i, nn, js, codes);
    double vo[G][D], xo[G][D], mo[G], deo[G];
    R C[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const int j = js[u];
      const double2 a = PV[4*(size_t)j], b = PV[4*(size_t)j + 1], c = PV[4*(size_t)j + 2], d = PV[4*(size_t)j + 3];
      vo[u][0] = a.x;
      if (D >= 2) vo[u][1 % D] = a.y;
      if (D >= 3) vo[u][2 % D] = b.x;
      deo[u] = b.y;
      mo[u] = c.x;
      xo[u][0] = c.y;
      if (D >= 2) xo[u][1 % D] = d.x;
      if (D >= 3) xo[u][2 % D] = d.y;
      C[u] = pdd[(size_t)(k0 + u)*ld + i];
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (k0 + u >= nn) continue;
      const int j = js[u], code = codes[u];
      const bool aFirst = code != 0 || i < j;
      if (code) image_pos<D>(g, code, xo[u]);
      // canonical: pacc_ab = (-m_b) C r_ab ; du_ab = (v_b^{1/2} - v_a^{1/2}) . pacc_ab
      const double mb = aFirst ? mo[u] : ms, ma = aFirst ? ms : mo[u];
      double duab = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const double dd = (double)(C[u]*(R)(xs[c] - xo[u][c]));     // the SPH pass's C*r, this lane's side
        const double v12o = refl_sign<D>(g, code, c)*vo[u][c];
        const double va = aFirst ? v12s[c] : v12o, vb = aFirst ? v12o : v12s[c];
        const double ddc = aFirst ? dd : -dd;
        duab += (vb - va)*((-mb)*ddc);
      }
      const double sgn = duab < 0.0 ? -1.0 : 1.0;
      const double wa = fmax(epsm, (aFirst ? de0s : deo[u])*sgn), wb = fmax(epsm, (aFirst ? deo[u] : de0s)*sgn);
      const double w = wa/(wa + wb);
      du += aFirst ? w*duab : (1.0 - w)*duab*ma/mb;
    }
  }
  return du;
}

// ---------------------------------------------------------------------------
// Full step from the start-of-step state with the midpoint derivatives
// (CheapSynchronousRK2.cc:106-109, SynchronousRK2.cc:115-118): the compatible energy
// partition (compatible_rate), then the increment and replace policies and the EOS.  v0 is
// read from f.v, the new velocity goes to f.v1.  The step variant also measures, per run, how far nodes have moved and h
// grown since the candidate list was built (see k_candidates).  ENERGY_ONLY: the
// package-level policy call (eps only).
// ---------------------------------------------------------------------------
template <int D, bool ENERGY_ONLY, typename R, int G>
__global__ void __launch_bounds__(kBlock) k_full_update(Fields f, Geo g, Phys ph, Scalars* sc, double multiplier,
                                                        int n) {
  const int i = xcd_block()*blockDim.x + threadIdx.x;
  const bool valid = i < n;
  int skr = 0;                       // candidate-list check (k_skin_check's measure, fused)
  double skd = 0.0, skg = 0.0;
  if (valid) {
  const double dt = ENERGY_ONLY ? multiplier : sc[f.run[i]].dt;
  const double du = compatible_rate<D, R, G>(f, g, i);
  const double eps1 = f.eps[i] + du*dt;
  f.eps[i] = eps1;
  if (ENERGY_ONLY) return;
  double x1[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    x1[a] = f.x[a][i] + dt*f.DxDt[a][i];
    f.x[a][i] = x1[a];
    double va = f.v[a][i] + dt*f.DvDt[a][i];
    if (ph.gravity) va = va + dt*f.DvDtG[a][i];
    f.v1[a][i] = va;
  }
  const double rho1 = fmin(ph.rho_max, fmax(ph.rho_min, f.rho[i] + dt*f.DrhoDt[i]));
  f.rho[i] = rho1;
  const double h1 = fmin(ph.hminInv, fmax(ph.hmaxInv, f.Hideal[i]));
  f.hinv[i] = h1;
  // end-of-step position and h^-1 for the grad-h pass that follows
  double2* PK = reinterpret_cast<double2*>(f.xh) + 2*(size_t)i;
  PK[0] = make_double2(x1[0], x1[1]);
  PK[1] = make_double2(x1[2], h1);
  // displacement and support growth since the run's candidate list was built
  skr = f.run[i];
  double d2 = 0.0;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double d = x1[a] - f.xb[a][i];
    if ((g.permask >> a) & 1) {
      const double L = g.xmax[a] - g.xmin[a];
      d = fabs(d);
      d = fmin(d, fabs(L - d));
    }
    d2 += d*d;
  }
  const double hb = f.hb[i];
  skd = sqrt(d2)*hb/ph.ext;
  skg = hb/h1;
  f.P[i] = ph.g1*rho1*eps1;
  f.cs[i] = sqrt(fmax(0.0, ph.gamma*ph.g1*eps1));
  }
  if (!ENERGY_ONLY) {
    const size_t stride = sizeof(Scalars)/sizeof(unsigned long long);
    run_max_bits(&sc[0].skinD, stride, skr, ord_bits(skd), valid);
    run_max_bits(&sc[0].skinG, stride, skr, ord_bits(skg), valid);
  }
}

// SynchronousRK2's trial half step (SynchronousRK2.cc:67-68: state.update with multiplier dt/2):
// the full policies, not timeAdvanceOnly -- the compatible energy partition of the start-of-step
// pair accelerations, the increments, H replaced by the ideal H (ReplaceBoundedState) and the
// EOS -- written to the midpoint arrays and position record; the start-of-step state stays.
template <int D, typename R, int G>
__global__ void __launch_bounds__(kBlock) k_half_update(Fields f, Geo g, Phys ph, const Scalars* sc, int n) {
  const int i = xcd_block()*blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double hdt = sc[f.run[i]].hdt;
  const double du = compatible_rate<D, R, G>(f, g, i);
  const double eps = f.eps[i] + du*hdt;
  double xm[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    xm[a] = f.x[a][i] + hdt*f.DxDt[a][i];
    f.xm[a][i] = xm[a];
    double va = f.v[a][i] + hdt*f.DvDt[a][i];
    if (ph.gravity) va = va + hdt*f.DvDtG[a][i];
    f.vm[a][i] = va;
  }
  const double rho = fmin(ph.rho_max, fmax(ph.rho_min, f.rho[i] + hdt*f.DrhoDt[i]));
  const double h = fmin(ph.hminInv, fmax(ph.hmaxInv, f.Hideal[i]));
  f.rhom[i] = rho;
  f.epsm[i] = eps;
  f.hinvm[i] = h;
  f.Pm[i] = ph.g1*rho*eps;
  f.csm[i] = sqrt(fmax(0.0, ph.gamma*ph.g1*eps));
  double2* P = reinterpret_cast<double2*>(f.xhm) + 2*(size_t)i;
  P[0] = make_double2(xm[0], xm[1]);
  P[1] = make_double2(xm[2], h);
}

// Half-step velocity of the compatible-energy partition, v^{1/2} = v + a*multiplier/2 with the
// evaluation's hydro acceleration (SpecificThermalEnergyPolicy.cc:97-99), packed with DepsDt and
// m (Fields::pkv).  sc == nullptr: hdt is the fixed multiplier/2; otherwise hdt is the fraction
// of each run's dt (1/4 for SynchronousRK2's trial half step, whose multiplier is dt/2).
template <int D>
__global__ void __launch_bounds__(kBlock) k_v12(Fields f, const Scalars* sc, double hdtScale, int n) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double hdt = sc ? hdtScale*sc[f.run[i]].dt : hdtScale;
  double w[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < D; ++a) w[a] = f.v[a][i] + f.DvDt[a][i]*hdt;
  double2* P = reinterpret_cast<double2*>(f.pkv) + 4*(size_t)i;
  P[0] = make_double2(w[0], w[1]);
  P[1] = make_double2(w[2], f.DepsDt[i]);
}


}  // namespace spx
