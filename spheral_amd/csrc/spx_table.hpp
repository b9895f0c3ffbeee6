// spx_table.hpp -- host-side TableKernel construction for the engine (FP64).
//
// Builds the per-interval parabola coefficients of W and dW/deta and the
// Wsum -> nPerh cubic-Hermite lookup exactly as the reference constructs its
// TableKernel (Kernel/TableKernel.cc:169-209):
//   * base kernels: Kernel/BSplineKernelInline.hh:8-91, Kernel/WendlandC4KernelInline.hh:8-95,
//     Kernel/NBSplineKernel.cc:20-110 (normalisation: Kernel/VolumeIntegrationFunctions.cc,
//     Utilities/simpsonsIntegration.hh)
//   * QuadraticInterpolator: Utilities/QuadraticInterpolator.cc:59-100 (odd sample count,
//     3-point exact parabola per interval), QuadraticInterpolatorInline.hh:19-35
//   * nPerh lookups: Kernel/TableKernel.cc:23-66 (sumKernelValues), CubicHermiteInterpolator
//     (Utilities/CubicHermiteInterpolator.cc: tridiagonal gradient knots, makeMonotonic),
//     Utilities/bisectRoot.hh.
// The device consumes the coefficient arrays as built here, so GPU kernel values
// are the same parabolas the host evaluates.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <vector>

namespace spx {

// NBSplineKernel (Kernel/NBSplineKernel.cc:20-110): centred B-spline with k = order + 1 knots as
// the truncated-power sum  sum_{i=0..k} (-1)^i C(k,i) (e - i + k/2)_+^p / p!  (p = k-1 for the
// value, k-2 for the derivative), in the reference's operation order.
inline int nb_factorial(int n) { int r = 1; for (int i = 1; i < n + 1; ++i) r *= i; return r; }
inline double nb_sum(int order, double e, int p) {
  const int k = order + 1;
  const double halfk = 0.5*k;
  double acc = 0.0;
  for (int i = 0; i <= k; ++i) {
    const double u = e - i + halfk;
    acc += std::pow(-1.0, (double)i)*(nb_factorial(k)/(nb_factorial(k - i)*nb_factorial(i)))*
           (u >= 0.0 ? std::pow(u, (double)p) : 0.0);
  }
  return acc;
}

struct KernelShape {
  int kind, dim, order;
  double norm, extent;
  KernelShape(int kind_, int dim_, int order_ = 5) : kind(kind_), dim(dim_), order(order_) {
    if (kind == 0) {  // BSpline
      extent = 2.0;
      norm = dim == 1 ? 2.0/3.0 : (dim == 2 ? 10.0/(7.0*M_PI) : 1.0/M_PI);
    } else if (kind == 2) {  // NBSpline: extent (order+1)/2 in integers, A = 1/Simpson volume integral
      extent = (double)((order + 1)/2);
      norm = 1.0;
      const unsigned bins = 10000u;
      const double dx = extent/bins;
      double acc = 0.0;
      for (unsigned i = 0u; i < bins + 1u; ++i) {
        const double r = 0.0 + i*dx;
        const double y = (dim == 1 ? 2.0 : dim == 2 ? 2.0*M_PI*r : 4.0*M_PI*r*r)*value(r);
        acc += (i == 0u || i == bins) ? y : (i % 2u == 0u ? 2.0*y : 4.0*y);
      }
      acc *= dx/3.0;
      norm = 1.0/acc;
    } else {          // WendlandC4
      extent = 1.0;
      norm = dim == 1 ? 1.5 : (dim == 2 ? 9.0/M_PI : 495.0/(32.0*M_PI));
    }
  }
  // Value, first and second derivative at Hdet = 1, written in the base kernels' own
  // arithmetic order so the tabulated samples are the same doubles.
  double value(double e) const {
    if (kind == 2) return e >= extent ? 0.0 : nb_sum(order, e, order)*(norm*1.0/nb_factorial(order));
    if (kind == 0) {
      if (e < 1.0) { const double e2 = e*e; return norm*1.0*(1.0 - 1.5*e2 + 0.75*e2*e); }
      if (e < 2.0) { const double t = 2.0 - e; return norm*1.0*0.25*(t*t*t); }
      return 0.0;
    }
    const double e2 = e*e;
    if (dim == 1) return norm*1.0*(std::pow(1.0 - e, 5)*(1.0 + 5.0*e + 8.0*e2))*(e < 1.0);
    return norm*1.0*(std::pow(1.0 - e, 6)*(1.0 + 6.0*e + (35.0/3.0)*e2))*(e < 1.0);
  }
  double grad(double e) const {
    if (kind == 2) return e >= extent ? 0.0 : nb_sum(order, e, order - 1)*(norm*1.0/nb_factorial(order - 1));
    if (kind == 0) {
      if (e < 1.0) return -norm*1.0*(3.0 - 2.25*e)*e;
      if (e < 2.0) { const double t = 2.0 - e; return -norm*1.0*0.75*(t*t); }
      return 0.0;
    }
    if (dim == 1) return norm*1.0*(-14.0*std::pow(1.0 - e, 4)*e*(1.0 + 4.0*e))*(e < 1.0);
    return norm*1.0*((56.0/3.0)*std::pow(e - 1.0, 5)*e*(5.0*e + 1.0))*(e < 1.0);
  }
};

// y = c0 + (c1 + c2 x) x on each of N1+1 intervals of [xmin, xmax].
struct Parabolas {
  std::size_t N1 = 0;
  double xmin = 0.0, xmax = 0.0, xstep = 0.0;
  std::vector<double> coef;

  template <typename F>
  void fit(double lo, double hi, std::size_t nsamp, const F& f) {
    if (nsamp % 2 == 0) ++nsamp;                        // both ends sampled
    const double dx = (hi - lo)/(nsamp - 1u);
    std::vector<double> y(nsamp);
    for (std::size_t i = 0; i < nsamp; ++i) y[i] = f(lo + i*dx);
    N1 = (nsamp - 1u)/2u - 1u;
    xstep = (hi - lo)/(N1 + 1u);
    coef.assign(3*(N1 + 1u), 0.0);
    for (std::size_t k = 0; k <= N1; ++k) {
      const double p0 = lo + k*xstep, p1 = p0 + 0.5*xstep, p2 = p0 + xstep;
      const double V[3][3] = {{1.0, p0, p0*p0}, {1.0, p1, p1*p1}, {1.0, p2, p2*p2}};
      // inverse(V) = adj(V) * (1/det V), det from the first-column cofactors
      double cf[3][3];
      for (int r = 0; r < 3; ++r)
        for (int s = 0; s < 3; ++s) {
          const int ra = (r + 1) % 3, rb = (r + 2) % 3, sa = (s + 1) % 3, sb = (s + 2) % 3;
          cf[r][s] = V[ra][sa]*V[rb][sb] - V[ra][sb]*V[rb][sa];
        }
      const double dV = cf[0][0]*V[0][0] + cf[1][0]*V[1][0] + cf[2][0]*V[2][0];
      const double rdet = 1.0/dV;
      const double b[3] = {y[2*k], y[2*k + 1], y[2*k + 2]};
      for (int r = 0; r < 3; ++r) {
        const double i0 = cf[0][r]*rdet, i1 = cf[1][r]*rdet, i2 = cf[2][r]*rdet;
        coef[3*k + r] = i0*b[0] + i1*b[1] + i2*b[2];
      }
    }
    xmin = lo;
    xmax = hi;
  }
  std::size_t base(double x) const { return 3u*std::min(N1, std::size_t(std::max(0.0, x - xmin)/xstep)); }
  double eval(double x, std::size_t b) const { return coef[b] + (coef[b + 1] + coef[b + 2]*x)*x; }
  double operator()(double x) const { return eval(x, base(x)); }
};

// Cubic Hermite on a uniform grid: vals[0,N) then knot gradients [N,2N).
struct HermiteTable {
  std::size_t N = 0;
  double xmin = 0.0, xmax = 0.0, xstep = 0.0;
  std::vector<double> vals;

  template <typename F>
  void fit(double lo, double hi, std::size_t n, const F& f) {
    N = n; xmin = lo; xmax = hi; xstep = (hi - lo)/(n - 1u);
    vals.assign(2*n, 0.0);
    for (std::size_t i = 0; i < n; ++i) vals[i] = f(lo + i*xstep);
    knots();
  }
  void knots() {
    // rows: [4,-1 | b0], [-0.5,4,-0.5 | 1.5 (y[k+1]-y[k-1])/dx], [-1,4 | bN]
    std::vector<double> lo(N, 0.0), di(N, 4.0), up(N, 0.0), rhs(N, 0.0);
    up[0] = -1.0; rhs[0] = 3.0*(vals[1] - vals[0])/xstep;
    lo[N - 1] = -1.0; rhs[N - 1] = 3.0*(vals[N - 1] - vals[N - 2])/xstep;
    for (std::size_t k = 1; k + 1 < N; ++k) { lo[k] = -0.5; up[k] = -0.5; rhs[k] = 1.5*(vals[k + 1] - vals[k - 1])/xstep; }
    for (std::size_t k = 1; k < N; ++k) { const double w = lo[k]/di[k - 1]; di[k] -= w*up[k - 1]; rhs[k] -= w*rhs[k - 1]; }
    vals[2*N - 1] = rhs[N - 1]/di[N - 1];
    for (std::size_t k = N - 1; k-- > 0;) vals[N + k] = (rhs[k] - up[k]*vals[N + k + 1])/di[k];
  }
  void monotone() {
    std::vector<double> sl(N - 1);
    const double r = 1.0/xstep;
    for (std::size_t k = 0; k + 1 < N; ++k) sl[k] = (vals[k + 1] - vals[k])*r;
    double* g = vals.data() + N;
    for (std::size_t k = 1; k + 1 < N; ++k) {
      if (sl[k - 1]*sl[k] <= 0.0) g[k] = 0.0;
      if (sl[k] == 0.0) { g[k] = 0.0; g[k + 1] = 0.0; }
    }
    for (bool again = true; again;) {
      again = false;
      for (std::size_t k = 0; k + 1 < N; ++k) {
        double al = g[k]/sl[k], be = g[k + 1]/sl[k];
        if (al < 0.0) { g[k] = 0.0; al = 0.0; again = true; }
        if (be < 0.0) { g[k + 1] = 0.0; be = 0.0; again = true; }
        const double tau = 3.0/std::sqrt(al*al + be*be);
        if (tau < 1.0) { g[k] = 0.99*tau*al*sl[k]; g[k + 1] = 0.99*tau*be*sl[k]; again = true; }
      }
    }
  }
  double operator()(double x) const {
    if (x < xmin) return vals[0] + vals[N]*(x - xmin);
    if (x > xmax) return vals[N - 1] + vals[2*N - 1]*(x - xmin);
    const std::size_t i0 = std::min(N - 2u, std::size_t(std::max(0.0, x - xmin)/xstep));
    const double t = std::max(0.0, std::min(1.0, (x - xmin - i0*xstep)/xstep));
    const double t2 = t*t, t3 = t*t2;
    return ((2.0*t3 - 3.0*t2 + 1.0)*vals[i0] + (-2.0*t3 + 3.0*t2)*vals[i0 + 1] +
            xstep*((t3 - 2.0*t2 + t)*vals[N + i0] + (t3 - t2)*vals[N + i0 + 1]));
  }
};

template <typename F>
double bisect(const F& f, double a, double b, double xacc = 1.0e-15, double yacc = 1.0e-10, unsigned maxit = 100u) {
  auto near0 = [&](double v) { return std::abs(v) <= yacc*std::max(1.0, std::abs(v)); };
  const double fa = f(a), fb = f(b);
  if (near0(fa)) return a;
  if (near0(fb)) return b;
  double x0 = fa < 0.0 ? a : b, x1 = fa < 0.0 ? b : a;
  for (unsigned it = 0; it < maxit; ++it) {
    const double dx = 0.5*(x1 - x0), mid = x0 + dx;
    if (std::abs(dx) <= xacc) return mid;
    const double fm = f(mid);
    if (near0(fm)) return mid;
    if (fm < 0.0) x0 = mid; else x1 = mid;
  }
  return 0.0;
}

struct KernelTable {
  int dim = 3;
  double extent = 0.0, minNperh = 0.0, maxNperh = 0.0;
  Parabolas W, gW;
  HermiteTable wsum, nperh;

  double Wsph(double e) const { return e < extent ? std::abs(gW(e)) : 0.0; }
  double wsumOf(double np) const {
    const double d = 1.0/np;
    double acc = 0.0, e = d;
    if (dim == 1) { for (; e < extent; e += d) acc += 2.0*Wsph(e); return acc; }
    if (dim == 2) { for (; e < extent; e += d) acc += 2.0*M_PI*e/d*Wsph(e); return std::sqrt(acc); }
    for (; e < extent; e += d) { const double q = e/d; acc += 4.0*M_PI*(q*q)*Wsph(e); }
    return std::pow(acc, 1.0/3.0);
  }
  void build(int kind, int dim_, unsigned numPoints, double minN, double maxN, int order = 5) {
    dim = dim_;
    const KernelShape k(kind, dim, order);
    W.fit(0.0, k.extent, numPoints, [&](double x) { return k.value(x); });
    gW.fit(0.0, k.extent, numPoints, [&](double x) { return k.grad(x); });
    extent = k.extent;
    maxNperh = maxN;
    minNperh = std::max(minN, 1.1/k.extent);
    if (maxNperh <= minNperh) maxNperh = 4.0*minNperh;
    wsum.fit(minNperh, maxNperh, numPoints, [&](double x) { return wsumOf(x); });
    const double w0 = wsum(minNperh), w1 = wsum(maxNperh);
    nperh.fit(w0, w1, numPoints, [&](double ws) {
      return bisect([&](double np) { return wsum(np) - ws; }, minNperh, maxNperh); });
    wsum.monotone();
    nperh.monotone();
  }
  void kernelAndGrad(double eta, double Hdet, double& w, double& gw) const {
    if (eta < extent) { const std::size_t b = W.base(eta); w = Hdet*W.eval(eta, b); gw = Hdet*gW.eval(eta, b); }
    else { w = 0.0; gw = 0.0; }
  }
  double nodesPerh(double ws) const { return std::max(0.0, nperh(ws)); }
};

}  // namespace spx
