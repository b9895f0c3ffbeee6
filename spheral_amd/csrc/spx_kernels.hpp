// spx_kernels.hpp -- neighbour build and the light pair passes of the SPH hot path.
//
// Every pair pass is a gather: one lane per target particle walks its ELLPACK
// neighbour list and accumulates only its own sums, so no atomics touch the
// conserved quantities.  Where a pair term must be bitwise antisymmetric (the
// pairwise acceleration and the compatible-energy partition) both lanes evaluate
// the pair in one canonical (a, b) order: an image partner makes the gathering
// node "a" (the reference's internal-ghost pair), otherwise the lower slot is "a".
//
// A batch holds independent runs of one problem type; particle i belongs to run
// f.run[i] with its own cell grid (RunGrid) and time step (Scalars).  Pairs never
// cross runs.  R is the arithmetic type of the pair math (double, or float in the
// mixed-precision mode where positions are differenced in double first).
#pragma once

#include "spx_device.hpp"

namespace spx {

// ---------------------------------------------------------------------------
// Neighbour image of a slot: position and velocity transforms (PeriodicBoundary
// copies, ReflectingBoundary::reflectOperator flips the normal component:
// Boundary/ReflectingBoundary.cc:183-197).
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

template <int D>
__device__ __forceinline__ int cell_coord(const RunGrid& gr, int a, double x) {
  int c = (int)floor((x - gr.origin[a])*gr.inv_cell[a]);
  return c < 0 ? 0 : (c >= gr.nc[a] ? gr.nc[a] - 1 : c);
}

// ---------------------------------------------------------------------------
// Boundary enforcement + sort keys.  Periodic violators are mapped through the
// opposite plane, reflecting violators mirrored with their velocity
// (Boundary/PlanarBoundary.cc:153-200, ReflectingBoundary.cc:465-473).  Keys order
// the particles of active runs by (run, cell Morton code), then the finished runs
// (kept whole, behind the active ones), then removed runs (largest key).  Active runs
// that keep their candidate list (build[run] == 0) get no Morton bits: the sort is
// stable, so they keep their order.
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(kBlock) k_keys(Fields f, Geo g, const RunGrid* grids, const int* rstate,
                                                 const int* rorder, const int* build, int n, int enforce, int withKeys,
                                                 int mortonBits, int runBits, uint64_t* keys, uint32_t* vals) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = f.run[i];
  uint64_t key = 0;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double xa = f.x[a][i];
    if (enforce) {
      for (int side = 1; side <= 2; ++side) {
        const int bc = side == 1 ? (((g.permask >> a) & 1) ? BC_PERIODIC : (((g.reflmask >> (2*a)) & 1) ? BC_REFLECTING : BC_NONE))
                                 : (((g.permask >> a) & 1) ? BC_PERIODIC : (((g.reflmask >> (2*a + 1)) & 1) ? BC_REFLECTING : BC_NONE));
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
    if (withKeys && build[r]) key |= spread_bits((uint64_t)cell_coord<D>(grids[r], a, xa), D) << a;
  }
  if (withKeys) {
    // layout: [rank | run | Morton], rank 0 active (runs by their place rorder), 1 done (by
    // slot); removed runs take the maximum key
    const int st = rstate[r];
    const int shift = mortonBits + runBits;
    keys[i] = st == RUN_ACTIVE ? (((uint64_t)rorder[r] << mortonBits) | key)
            : (st == RUN_DONE ? ((1ull << shift) | ((uint64_t)r << mortonBits)) : ~0ull);
    vals[i] = (uint32_t)i;
  }
}

// Gather a set of arrays through a permutation (dst[p] = src[perm[p]]).
struct GatherSet {
  const double* src[64];
  double* dst[64];
  int count;
};
__global__ void __launch_bounds__(kBlock) k_gather(GatherSet s, const uint32_t* perm, int n) {
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t q = perm[p];
  for (int k = 0; k < s.count; ++k) s.dst[k][p] = s.src[k][q];
}
// Values given in a run's own node order into the engine's current particle order: slot p
// holds node (oid[p] & low 40 bits) of its run.
__global__ void __launch_bounds__(kBlock) k_from_local(double* __restrict__ dst, const double* __restrict__ src,
                                                       const int64_t* __restrict__ oid, int n) {
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p < n) dst[p] = src[oid[p] & (((int64_t)1 << 40) - 1)];
}
__global__ void __launch_bounds__(kBlock) k_gather_ids(const int64_t* src, int64_t* dst, const int32_t* rsrc, int32_t* rdst,
                                                       const uint32_t* perm, int n) {
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p < n) { dst[p] = src[perm[p]]; rdst[p] = rsrc[perm[p]]; }
}

template <int D>
__device__ __forceinline__ int lin_cell(const RunGrid& g, const int* c) {
  int l = c[D - 1];
  for (int a = D - 2; a >= 0; --a) l = l*g.nc[a] + c[a];
  return g.cellBase + l;
}
template <int D>
__device__ __forceinline__ void unspread(uint64_t key, int mortonBits, int* c) {
  for (int a = 0; a < D; ++a) {
    uint64_t v = 0;
    for (int b = 0; b < 21 && D*b + a < mortonBits; ++b) v |= ((key >> (D*b + a)) & 1ull) << b;
    c[a] = (int)v;
  }
}

// Cell ranges of the sorted keys of the runs that build candidate lists: start[cell],
// end[cell] (both 0 when empty).  The low mortonBits of a key are the cell's Morton code;
// a run's particles are one contiguous segment (run[] tells where it ends).
// With cell levels, each level's table: a level-l cell is the key prefix above the low
// D*(levels-1-l) bits, one contiguous range of the sorted particles.
template <int D>
__device__ __forceinline__ int level_cells(const RunGrid& g, int l, int* ncl) {
  const int sh = g.levels - 1 - l;
  int t = 1;
  for (int a = 0; a < D; ++a) { ncl[a] = ((g.nc[a] - 1) >> sh) + 1; t *= ncl[a]; }
  return t;
}
template <int D>
__device__ __forceinline__ int lin_level_cell(const RunGrid& g, int l, const int* ncl, const int* c) {
  int q = c[D - 1];
  for (int a = D - 2; a >= 0; --a) q = q*ncl[a] + c[a];
  return g.levelBase[l] + q;
}
template <int D>
__global__ void __launch_bounds__(kBlock) k_cell_bounds(const uint64_t* keys, const int32_t* run, int n,
                                                        const RunGrid* grids, const int* build, int mortonBits,
                                                        int* start, int* end) {
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int r = run[p];
  if (!build[r]) return;
  const uint64_t mask = (1ull << mortonBits) - 1ull;
  const uint64_t k = keys[p] & mask;
  const RunGrid& gr = grids[r];
  const bool first = p == 0 || run[p - 1] != r, last = p == n - 1 || run[p + 1] != r;
  int c[3];
  if (gr.levels <= 1) {
    if (first || (keys[p - 1] & mask) != k) { unspread<D>(k, mortonBits, c); start[lin_cell<D>(gr, c)] = p; }
    if (last || (keys[p + 1] & mask) != k) { unspread<D>(k, mortonBits, c); end[lin_cell<D>(gr, c)] = p + 1; }
    return;
  }
  const uint64_t kp = first ? 0 : (keys[p - 1] & mask), kn = last ? 0 : (keys[p + 1] & mask);
  for (int l = 0; l < gr.levels; ++l) {
    const int sh = D*(gr.levels - 1 - l);
    const uint64_t q = k >> sh;
    int ncl[3] = {1, 1, 1};
    level_cells<D>(gr, l, ncl);
    if (first || (kp >> sh) != q) { unspread<D>(q, mortonBits, c); start[lin_level_cell<D>(gr, l, ncl, c)] = p; }
    if (last || (kn >> sh) != q) { unspread<D>(q, mortonBits, c); end[lin_level_cell<D>(gr, l, ncl, c)] = p + 1; }
  }
}

// Sort keys of the particles of the runs that rebuild (build[run] != 0), compacted: run r's
// particles go to [sub[3r+1], +n) in position order, keyed [its segment index | cell Morton
// code] (sub: 3 ints per run: segment start, compacted start, segment index).
__global__ void __launch_bounds__(kBlock) k_subset_keys(const uint64_t* keys, const int32_t* run, const int* build,
                                                        const int* sub, int mortonBits, int n, uint64_t* skeys,
                                                        uint32_t* spos) {
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int r = run[p];
  if (!build[r]) return;
  const int k = sub[3*r + 1] + (p - sub[3*r]);
  const uint64_t mask = (1ull << mortonBits) - 1ull;
  skeys[k] = ((uint64_t)sub[3*r + 2] << mortonBits) | (keys[p] & mask);
  spos[k] = (uint32_t)p;
}

// Gather over a list of positions: dst[a][pos[k]] = src[a][from ? from[k] : pos[k]], oid alike;
// with sortedKeys, keysOut[pos[k]] = sortedKeys[k].
__global__ void __launch_bounds__(kBlock) k_gather_list(GatherSet s, const uint32_t* pos, const uint32_t* from,
                                                        const int64_t* oidSrc, int64_t* oidDst,
                                                        const uint64_t* sortedKeys, uint64_t* keysOut, int nsub) {
  const int k = blockIdx.x*blockDim.x + threadIdx.x;
  if (k >= nsub) return;
  const uint32_t p = pos[k], q = from ? from[k] : p;
  for (int a = 0; a < s.count; ++a) s.dst[a][p] = s.src[a][q];
  oidDst[p] = oidSrc[q];
  if (sortedKeys) keysOut[p] = sortedKeys[k];
}

// Running per-run minimum of a positive double: one atomic per wavefront when its lanes share a run.
__device__ __forceinline__ void run_min(Scalars* sc, int r, double v, bool valid) {
  const int r0 = __shfl(r, 0);
  const bool uniform = __all(!valid || r == r0);
  if (uniform) {
    double m = valid ? v : 1.7976931348623157e308;
    for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off));
    // most waves do not lower the running minimum: read first, then update atomically
    const unsigned long long b = ord_bits(m);
    if ((threadIdx.x & 63) == 0 && __any(valid) && b < sc[r0].dtBits) atomicMin(&sc[r0].dtBits, b);
  } else if (valid) {
    const unsigned long long b = ord_bits(v);
    if (b < sc[r].dtBits) atomicMin(&sc[r].dtBits, b);
  }
}

// Packed neighbour records (Fields::xh, Fields::vmr) of the current state, and nodeScale =
// 1/(lambda_max(H) nPerh) of each node (GenericHydro::dt, Physics/GenericHydro.cc:142-150).
template <int D>
__global__ void __launch_bounds__(kBlock) k_pack(Fields f, Phys ph, int n) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double h = f.hinv[i];
  double2* P = reinterpret_cast<double2*>(f.xh) + 2*(size_t)i;
  P[0] = make_double2(f.x[0][i], D >= 2 ? f.x[1 % D][i] : 0.0);
  P[1] = make_double2(D >= 3 ? f.x[2 % D][i] : 0.0, h);
  double2* V = reinterpret_cast<double2*>(f.vmr) + 2*(size_t)i;
  V[0] = make_double2(f.v[0][i], D >= 2 ? f.v[1 % D][i] : 0.0);
  V[1] = make_double2(D >= 3 ? f.v[2 % D][i] : 0.0, f.m[i]);
  f.ns[i] = 1.0/max_eigen_iso<D>(h)/ph.nperh;
}

// Position and h^-1 of node j from a 4-double (x, h^-1) record array.
template <int D>
__device__ __forceinline__ void load_xh(const double* xh, int j, double* x, double& h) {
  const double2* P = reinterpret_cast<const double2*>(xh) + 2*(size_t)j;
  const double2 a = P[0], b = P[1];
  x[0] = a.x;
  if (D >= 2) x[1 % D] = a.y;
  if (D >= 3) x[2 % D] = b.x;
  h = b.y;
}
// Velocity and mass of node j from a 4-double (v, m) record array.
template <int D>
__device__ __forceinline__ void load_vm(const double* vm, int j, double* v, double& m) {
  const double2* P = reinterpret_cast<const double2*>(vm) + 2*(size_t)j;
  const double2 a = P[0], b = P[1];
  v[0] = a.x;
  if (D >= 2) v[1 % D] = a.y;
  if (D >= 3) v[2 % D] = b.x;
  m = b.y;
}

// Store the 8-entry chunk holding entry k of slot i (two 16-byte stores).
__device__ __forceinline__ void flush_chunk(uint32_t* nbr, int k, int i, size_t ld, const uint32_t* c) {
  uint4* dst = reinterpret_cast<uint4*>(nbr + ell_at(k & ~7, i, ld));
  dst[0] = make_uint4(c[0], c[1], c[2], c[3]);
  dst[1] = make_uint4(c[4], c[5], c[6], c[7]);
}

// ---------------------------------------------------------------------------
// ConnectivityMap::computeConnectivity (Neighbor/ConnectivityMap.cc:905-947): j is a
// neighbour of i iff |Hi rij|^2 <= ext^2 or |Hj rij|^2 <= ext^2, i != j (an image of i
// itself counts, as the reference's ghost of i does), evaluated with the reference's
// rounding (no contraction) so the sets match exactly.
//
// Two stages.  k_candidates builds, for the runs whose list is stale, a candidate list
// with the criterion widened to ext (1 + skin): one lane per particle walks the 3^D cells
// (of side >= the widened support) around its own; lanes of a wavefront mostly share a
// cell, so candidate reads are wave-uniform.  It records the positions and h^-1 the list
// was built from.  k_neighbor_filter applies the exact criterion to the candidates every
// step.  The candidates stay complete while no pair can have closed in from beyond the
// widened radius: with D = max |x - xb| hb/ext and G = max hb/h^-1 over the run's nodes,
// |rij(now)| <= max(Si, Sj)(now) implies |rij(built)| <= max(Si, Sj)(built) (G + 2D), so a
// list is valid while G + 2D <= 1 + skin (the full update measures both after every step).
// Reflecting images are isometries of the node, and periodic images are re-chosen as the
// nearest copy each step (the box spans >= 3 widened supports), so both obey the bound.
// ---------------------------------------------------------------------------
// Image code of the same pair seen from the other node: a periodic image flips side, a
// reflecting image is its own inverse.
__device__ __forceinline__ int inverse_image(const Geo& g, int code) {
  int out = code;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int c = (code >> (2*a)) & 3;
    if (c && ((g.permask >> a) & 1)) out = (out & ~(3 << (2*a))) | ((3 - c) << (2*a));
  }
  return out;
}

// Candidate lists of the runs that rebuild: each node gathers the j within its own widened
// support (|Hi rij|^2 <= ext^2 (1 + skin)^2); see k_candidates for the cells it scans.
//
// The rebuilding runs' segments, listed as (first position, first index in the compacted
// range [0, nBuild)) in increasing order of both: compacted index k -> particle position, and
// (with seg_of) the segment that holds it.
__device__ __forceinline__ int seg_of(const int2* seg, int nseg, int k) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg[mid].y <= k) lo = mid; else hi = mid - 1;
  }
  return lo;
}
__device__ __forceinline__ int build_pos(const int2* seg, int nseg, int k) {
  const int s = seg_of(seg, nseg, k);
  return seg[s].x + (k - seg[s].y);
}

// Launched over the rebuilding runs' particles only (seg, nBuild): a step that rebuilds a few
// small runs of a large batch neither walks the others nor leaves the work to the one XCD whose
// index range holds those runs.
//
// Cells.  A run whose h is near uniform has one grid of cells at least as wide as the widest
// widened support, and each node scans the 3^D cells around its own (image cells through
// periodic and reflecting planes).  Where h varies by 2x or more:
//  * open boundaries: levels.  Level-l cells are at least support/2^(l+2) wide (l = 0..levels-1;
//    the grid is that of the finest level, level-l cells are 2^(levels-1-l) finest cells per
//    axis, so each is one contiguous range of the Morton-sorted particles, with a table per
//    level).  A node works at the finest level l with support/2^l >= its widened support, whose
//    cells are at least a quarter of that support, and scans the 3 to 9 cells per axis the
//    support reaches: dense regions scan cells of their own size, sparse ones cells of theirs.
//  * periodic or reflecting axes: one grid, cells shrunk toward the smaller supports (by up to
//    4x while the grid stays within ~8 cells per node), and each node scans the cells its own
//    support reaches.
// Every node collects each j within its own widened support; where j's support does not reach
// i, i hands j the mirrored entry (tlist, merged by k_candidates_merge).  Together each list
// holds exactly the j with |Hi rij| or |Hj rij| within the widened extent.
template <int D>
__global__ void __launch_bounds__(kBlock) k_candidates(Fields f, Geo g, const RunGrid* grids, const int2* seg, int nseg,
                                                       int nBuild, const int* cstart, const int* cend, uint32_t* tlist,
                                                       int* tcnt, int tcap, int* overflow) {
  const int k = xcd_block()*blockDim.x + threadIdx.x;
  if (k >= nBuild) return;
  const int sg = seg_of(seg, nseg, k);
  const int i = seg[sg].x + (k - seg[sg].y);
  const int r = f.run[i];
  const RunGrid gr = grids[r];
  double xi[D];
  int ci[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    xi[a] = f.x[a][i];
    ci[a] = cell_coord<D>(gr, a, xi[a]);
    f.xb[a][i] = xi[a];
  }
  const double hi = f.hinv[i];
  f.hb[i] = hi;
  const double ext2c = gr.ext2c;
  const double reach = sqrt(ext2c)/hi;
  // FP32 distance pre-filter with a 1e-4 relative margin (far above its own rounding): only the
  // FP64 criterion decides membership
  const float reach2 = (float)(ext2c/(hi*hi)*(1.0 + 1.0e-4));
  const size_t ld = (size_t)f.ld;
  const int cap = f.ccap;
  int count = 0;
  uint32_t chunk[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  auto visit = [&](int j, int code) {
    if (j == i && code == 0) return;
    double xj[D];
#pragma unroll
    for (int a = 0; a < D; ++a) xj[a] = f.x[a][j];
    if (code) image_pos<D>(g, code, xj);
    float r2 = 0.0f;
#pragma unroll
    for (int a = 0; a < D; ++a) { const float d = (float)(xi[a] - xj[a]); r2 += d*d; }
    if (r2 > reach2) return;
    const double hj = f.hinv[j];
    double e2i = 0.0, e2j = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const double rr = dsub(xi[a], xj[a]);
      const double ei = dmul(hi, rr), ej = dmul(hj, rr);
      e2i = a == 0 ? dmul(ei, ei) : dadd(e2i, dmul(ei, ei));
      e2j = a == 0 ? dmul(ej, ej) : dadd(e2j, dmul(ej, ej));
    }
    if (!(e2i <= ext2c)) return;
    // write-combine 8 entries in registers, flush them as one 32-byte chunk
    const uint32_t e = (uint32_t)j | ((uint32_t)code << kImgShift);
    const int s = count & 7;
#pragma unroll
    for (int u = 0; u < 8; ++u) chunk[u] = s == u ? e : chunk[u];
    ++count;
    if (s == 7 && count <= cap) flush_chunk(f.cand, count - 1, i, ld, chunk);
    if (!(e2j <= ext2c)) {
      // j's own support does not reach i: j receives the pair from here
      const int slot = atomicAdd(&tcnt[j], 1);
      if (slot < tcap) tlist[ell_at(slot, j, ld)] = (uint32_t)i | ((uint32_t)inverse_image(g, code) << kImgShift);
    }
  };
  if (gr.levels > 1) {
    // the node's level: the finest l with support/2^l >= its reach; those cells are at least
    // support/2^(l+2) wide, a quarter of the reach or more, so it scans 3 to 9 cells per axis
    int lv = 0;
    double w = 0.5*gr.support;
    while (lv + 1 < gr.levels && w >= reach) { ++lv; w *= 0.5; }
    const int sh = gr.levels - 1 - lv;
    int cl[3] = {0, 0, 0}, ncl[3] = {1, 1, 1}, kl[3] = {0, 0, 0};
    level_cells<D>(gr, lv, ncl);
#pragma unroll
    for (int a = 0; a < D; ++a) {
      cl[a] = ci[a] >> sh;
      kl[a] = max(1, (int)ceil(reach*gr.inv_cell[a]/(double)(1 << sh)*(1.0 + 1.0e-12)));
    }
    const int kz = D >= 3 ? kl[2] : 0, ky = D >= 2 ? kl[1] : 0;
    for (int dz = -kz; dz <= kz; ++dz)
      for (int dy = -ky; dy <= ky; ++dy)
        for (int dx = -kl[0]; dx <= kl[0]; ++dx) {
          const int off[3] = {dx, dy, dz};
          int cc[3] = {0, 0, 0};
          bool skip = false;
#pragma unroll
          for (int a = 0; a < D; ++a) {
            cc[a] = cl[a] + off[a];
            skip |= cc[a] < 0 || cc[a] >= ncl[a];
          }
          if (skip) continue;
          const int cell = lin_level_cell<D>(gr, lv, ncl, cc);
          const int q1 = cend[cell];
          for (int j = cstart[cell]; j < q1; ++j) visit(j, 0);
        }
  } else {
    // stencil half-width per axis: the cells the node's own widened support reaches
    int kw[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < D; ++a) kw[a] = max(1, (int)ceil(reach*gr.inv_cell[a]*(1.0 + 1.0e-12)));
    const int kz = D >= 3 ? kw[2] : 0, ky = D >= 2 ? kw[1] : 0;
    for (int dz = -kz; dz <= kz; ++dz)
      for (int dy = -ky; dy <= ky; ++dy)
        for (int dx = -kw[0]; dx <= kw[0]; ++dx) {
          const int off[3] = {dx, dy, dz};
          int cc[3] = {0, 0, 0};
          int code = 0;
          bool skip = false;
#pragma unroll
          for (int a = 0; a < D; ++a) {
            int c = ci[a] + off[a];
            const bool per = (g.permask >> a) & 1;
            if (c < 0) {
              if (per) { c += gr.nc[a]; code |= 1 << (2*a); }
              else if ((g.reflmask >> (2*a)) & 1) { c = -c - 1; code |= 1 << (2*a); skip |= c >= gr.nc[a]; }
              else skip = true;
            } else if (c >= gr.nc[a]) {
              if (per) { c -= gr.nc[a]; code |= 2 << (2*a); }
              else if ((g.reflmask >> (2*a + 1)) & 1) { c = 2*gr.nc[a] - 1 - c; code |= 2 << (2*a); skip |= c < 0; }
              else skip = true;
            }
            cc[a] = c;
          }
          if (skip) continue;
          const int cell = lin_cell<D>(gr, cc);
          const int q1 = cend[cell];
          for (int j = cstart[cell]; j < q1; ++j) visit(j, code);
        }
  }
  if ((count & 7) && count <= cap) flush_chunk(f.cand, count - 1, i, ld, chunk);
  f.ccnt[i] = count;
  if (count > cap) atomicMax(overflow, count);
}

// Appends each rebuilding node's handed-over entries (tlist) to its candidate list, in
// ascending entry order (the order they arrived in is not deterministic).  Over the rebuilding
// runs' particles (seg, nBuild); each wavefront takes its 64 nodes in turn, only those that
// received entries: lane l ranks entries l, l+64, ...
// against all the node's entries, read 64 at a time and broadcast lane by lane
// (v_readlane); entry k goes to c0 + (number of smaller entries), entries being distinct.
// Batches whose h is uniform hand nothing over, so most waves only read their counts.
// overflow[0]: the candidate capacity needed, overflow[1]: the handed-over capacity needed.
__global__ void __launch_bounds__(kBlock) k_candidates_merge(Fields f, const int2* seg, int nseg, int nBuild,
                                                             const uint32_t* tlist, const int* tcnt, int tcap,
                                                             int* overflow) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x*blockDim.x + threadIdx.x;
  const int me = k < nBuild ? build_pos(seg, nseg, k) : 0;
  int t = 0, c0 = 0;
  bool ok = false;
  if (k < nBuild) {
    t = tcnt[me];
    if (t > tcap) {
      atomicMax(&overflow[1], t);
    } else if (t > 0) {
      c0 = f.ccnt[me];
      if (c0 + t > f.ccap) atomicMax(&overflow[0], c0 + t);
      else ok = true;
    }
  }
  const size_t ld = (size_t)f.ld;
  unsigned long long todo = __ballot(ok);
  while (todo) {
    const int l = __ffsll((long long)todo) - 1;
    todo &= todo - 1;
    const int i = __shfl(me, l), ti = __shfl(t, l), ci = __shfl(c0, l);
    for (int kb = 0; kb < ti; kb += 64) {
      const int k = kb + lane;
      const uint32_t e = k < ti ? tlist[ell_at(k, i, ld)] : 0u;
      int rank = 0;
      for (int cb = 0; cb < ti; cb += 64) {
        const int v = cb + lane < ti ? (int)tlist[ell_at(cb + lane, i, ld)] : -1;
        const int m = min(64, ti - cb);
        for (int q = 0; q < m; ++q) rank += (uint32_t)__builtin_amdgcn_readlane(v, q) < e ? 1 : 0;
      }
      if (k < ti) f.cand[ell_at(ci + rank, i, ld)] = e;
    }
  }
  if (ok) f.ccnt[me] = c0 + t;
}

// The exact criterion over each active node's candidates; writes the ELL neighbour list in
// candidate order.  Periodic image codes are re-chosen from the current positions (nearest
// copy); reflecting codes are kept.
template <int D, int G>
__global__ void __launch_bounds__(kBlock) k_neighbor_filter(Fields f, Geo g, Phys ph, int n, int* overflow) {
  const int i = xcd_block()*blockDim.x + threadIdx.x;
  if (i >= n) return;
  double xi[D];
#pragma unroll
  for (int a = 0; a < D; ++a) xi[a] = f.x[a][i];
  const double hi = f.hinv[i];
  const size_t ld = (size_t)f.ld;
  const int cap = f.cap;
  const int nc = f.ccnt[i];
  int count = 0;
  uint32_t chunk[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int k0 = 0; k0 < nc; k0 += G) {
    int js[G], codes[G];
    ell_group<G>(f.cand, ld, k0, i, nc, js, codes);
    double xj[G][D], hj[G];
#pragma unroll
    for (int u = 0; u < G; ++u) load_xh<D>(f.xh, js[u], xj[u], hj[u]);
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (k0 + u >= nc) continue;
      int code = codes[u];
#pragma unroll
      for (int a = 0; a < D; ++a) {
        if ((g.permask >> a) & 1) {
          const double half = 0.5*(g.xmax[a] - g.xmin[a]);
          const double d = xj[u][a] - xi[a];
          const int c = d > half ? 1 : (d < -half ? 2 : 0);
          code = (code & ~(3 << (2*a))) | (c << (2*a));
        }
      }
      if (js[u] == i && code == 0) continue;
      if (code) image_pos<D>(g, code, xj[u]);
      double e2i = 0.0, e2j = 0.0;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        const double rr = dsub(xi[a], xj[u][a]);
        const double ei = dmul(hi, rr), ej = dmul(hj[u], rr);
        e2i = a == 0 ? dmul(ei, ei) : dadd(e2i, dmul(ei, ei));
        e2j = a == 0 ? dmul(ej, ej) : dadd(e2j, dmul(ej, ej));
      }
      if (e2i <= ph.ext2 || e2j <= ph.ext2) {
        const uint32_t e = (uint32_t)js[u] | ((uint32_t)code << kImgShift);
        const int s = count & 7;
#pragma unroll
        for (int q = 0; q < 8; ++q) chunk[q] = s == q ? e : chunk[q];
        ++count;
        if (s == 7 && count <= cap) flush_chunk(f.nbr, count - 1, i, ld, chunk);
      }
    }
  }
  if ((count & 7) && count <= cap) flush_chunk(f.nbr, count - 1, i, ld, chunk);
  f.cnt[i] = count;
  if (count > cap) atomicMax(overflow, count);
}

// Per-run maximum of an order-preserving double key: a wavefront whose lanes share a run
// reduces first, and an atomic is issued only when it would raise the running maximum
// (most wavefronts do not: reading first keeps them off the contended word).
__device__ __forceinline__ void run_max_bits(unsigned long long* word, size_t stride, int r, unsigned long long b,
                                             bool valid) {
  const int r0 = __shfl(r, 0);
  if (__all(!valid || r == r0)) {
    unsigned long long m = valid ? b : 0ull;
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(m, off);
      m = o > m ? o : m;
    }
    unsigned long long* w = word + (size_t)r0*stride;
    if ((threadIdx.x & 63) == 0 && __any(valid) && m > __atomic_load_n(w, __ATOMIC_RELAXED)) atomicMax(w, m);
  } else if (valid) {
    unsigned long long* w = word + (size_t)r*stride;
    if (b > __atomic_load_n(w, __ATOMIC_RELAXED)) atomicMax(w, b);
  }
}

// Moves the candidate lists of runs whose particle segment shifted in a reorder (a run
// that keeps its list keeps its order, so old slot = new slot - delta):
// dst row p = src row p - delta with every index shifted by delta.  delta[run] ==
// kNoRemap marks runs whose rows are not moved.
constexpr int kNoRemap = INT32_MIN;
__global__ void __launch_bounds__(kBlock) k_remap_candidates(const uint32_t* src, const int32_t* srcCnt, uint32_t* dst,
                                                             int32_t* dstCnt, const int32_t* run, const int* delta,
                                                             size_t ld, int n) {
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int d = delta[run[p]];
  if (d == kNoRemap) return;
  const int q = p - d;
  const int cnt = srcCnt[q];
  dstCnt[p] = cnt;
  for (int c = 0; c < cnt; c += 8) {
    const uint4* s = reinterpret_cast<const uint4*>(src + ell_at(c, q, ld));
    uint4* t = reinterpret_cast<uint4*>(dst + ell_at(c, p, ld));
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      uint4 v = s[h];
      v.x += (uint32_t)d; v.y += (uint32_t)d; v.z += (uint32_t)d; v.w += (uint32_t)d;
      t[h] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// computeSPHSumMassDensity (SPH/computeSPHSumMassDensity.cc:35-85) in gather form
// -- node i receives m_j W(|Hj ri - Hj rj|, det Hj) -- fused with the GenericHydro::dt
// vote (Physics/GenericHydro.cc:112-381): per-node limits on the new density and the
// pairwise velocity-difference limit, min-reduced per run.
// FILT: the pass walks the candidate lists instead, applies the neighbour criterion of
// k_neighbor_filter to each candidate and writes the ELL neighbour list as it goes, so the
// connectivity filter and the first pair pass of the step read the neighbours once.
// ---------------------------------------------------------------------------
template <int D, bool DENS, bool DT, typename R, int G, bool FILT, typename CP>
__device__ __forceinline__ void density_dt_node(const Fields& f, const Geo& g, const Phys& ph, const DevTable& tab,
                                                CP coef, int n, Scalars* sc, int* overflow) {
  const int i = xcd_block()*blockDim.x + threadIdx.x;
  double dtmin = 1.7976931348623157e308;
  const bool valid = i < n;
  int r = 0;
  if (valid) {
    r = f.run[i];
    double xi[D], vi[D];
#pragma unroll
    for (int a = 0; a < D; ++a) { xi[a] = f.x[a][i]; vi[a] = f.v[a][i]; }
    const double nsi = f.ns[i];
    double rho = 0.0;
    if (DENS) {
      const R hi = (R)f.hinv[i];
      R w, gw, graw;
      kernel_wgw_c<R>(coef, tab, (R)0, hdet_iso<D, R>(hi), w, gw, graw);   // self term W(|Hi ri - Hi ri|) = W(0)
      rho = f.m[i]*(double)w;
    }
    // pairwise limit min(ns_i, ns_j)/|v_ij| (GenericHydro.cc:196-210).  For isotropic H the
    // nodeScale is 1/(h^-1 nPerh) (lambda_max(h^-1 I) = h^-1; the reference's eigen-solver may
    // differ from it in the last bits), so the largest |v_ij|^2/min(ns)^2 is tracked as
    // max |v_ij|^2 max(h_i^-1, h_j^-1)^2, scaled by nPerh^2 at the end: no per-pair division
    // and no nodeScale read per neighbour
    const double hiD = f.hinv[i];
    double vmaxh = 0.0;
    const size_t ld = (size_t)f.ld;
    const int nn = FILT ? f.ccnt[i] : f.cnt[i];
    const double hiF = FILT ? f.hinv[i] : 0.0;
    int count = 0;
    uint32_t chunk[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k0 = 0; k0 < nn; k0 += G) {
      int js[G], codes[G];
      ell_group<G>(FILT ? f.cand : f.nbr, ld, k0, i, nn, js, codes);
      double xj[G][D], vj[G][D], hj[G], mj[G];
      bool acc[G];
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const int j = js[u];
        load_xh<D>(f.xh, j, xj[u], hj[u]);
        if (!FILT) load_vm<D>(f.vmr, j, vj[u], mj[u]);
        acc[u] = !FILT && k0 + u < nn;
      }
      if (FILT) {
        // the exact neighbour criterion (k_neighbor_filter) on the group's candidates first;
        // velocities and masses are then read for the accepted ones only
#pragma unroll
        for (int u = 0; u < G; ++u) {
          if (k0 + u >= nn) continue;
          int code = codes[u];
#pragma unroll
          for (int a = 0; a < D; ++a) {
            if ((g.permask >> a) & 1) {
              const double half = 0.5*(g.xmax[a] - g.xmin[a]);
              const double d = xj[u][a] - xi[a];
              const int c = d > half ? 1 : (d < -half ? 2 : 0);
              code = (code & ~(3 << (2*a))) | (c << (2*a));
            }
          }
          codes[u] = code;
          if (js[u] == i && code == 0) continue;
          if (code) image_pos<D>(g, code, xj[u]);
          double e2i = 0.0, e2j = 0.0;
#pragma unroll
          for (int a = 0; a < D; ++a) {
            const double rr = dsub(xi[a], xj[u][a]);
            const double ei = dmul(hiF, rr), ej = dmul(hj[u], rr);
            e2i = a == 0 ? dmul(ei, ei) : dadd(e2i, dmul(ei, ei));
            e2j = a == 0 ? dmul(ej, ej) : dadd(e2j, dmul(ej, ej));
          }
          if (!(e2i <= ph.ext2 || e2j <= ph.ext2)) continue;
          acc[u] = true;
          const uint32_t e = (uint32_t)js[u] | ((uint32_t)code << kImgShift);
          const int sl = count & 7;
#pragma unroll
          for (int q = 0; q < 8; ++q) chunk[q] = sl == q ? e : chunk[q];
          ++count;
          if (sl == 7 && count <= f.cap) flush_chunk(f.nbr, count - 1, i, ld, chunk);
        }
#pragma unroll
        for (int u = 0; u < G; ++u)
          if (acc[u]) load_vm<D>(f.vmr, js[u], vj[u], mj[u]);
      }
#pragma unroll
      for (int u = 0; u < G; ++u) {
        if (!acc[u]) continue;
        const int code = codes[u];
        if (DENS) {
          if (!FILT && code) image_pos<D>(g, code, xj[u]);
          R s = (R)0;
          if (sizeof(R) == sizeof(double)) {
#pragma unroll
            for (int a = 0; a < D; ++a) { const double d = hj[u]*xi[a] - hj[u]*xj[u][a]; s += (R)(d*d); }
          } else {
#pragma unroll
            for (int a = 0; a < D; ++a) { const R d = (R)hj[u]*(R)(xi[a] - xj[u][a]); s += d*d; }
          }
          R w, gw, graw;
          kernel_wgw_c<R>(coef, tab, sqrt(s), hdet_iso<D, R>((R)hj[u]), w, gw, graw);
          rho += mj[u]*(double)w;
        }
        if (DT) {
          double s = 0.0;
#pragma unroll
          for (int a = 0; a < D; ++a) { const double d = vi[a] - refl_sign<D>(g, code, a)*vj[u][a]; s += d*d; }
          const double hm = fmax(hiD, hj[u]);
          vmaxh = fmax(vmaxh, s*(hm*hm));
        }
      }
    }
    if (FILT) {
      if ((count & 7) && count <= f.cap) flush_chunk(f.nbr, count - 1, i, ld, chunk);
      f.cnt[i] = count;
      if (count > f.cap) atomicMax(overflow, count);
    }
    if (DENS) f.rho[i] = rho; else rho = f.rho[i];
    if (DT) {
      const double tiny = 2.220446049250313e-16;
      // min(ns_i, ns_j)/max(tiny, |v_ij|) minimised over the pairs
      const double vmax2 = vmaxh*(ph.nperh*ph.nperh);
      if (vmax2 > 0.0) dtmin = fmin(dtmin, 1.0/sqrt(vmax2));
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
  if (DT) run_min(sc, r, dtmin, valid);
}

// The kernel is a thin wrapper around the node body: compiled this way the filtering pass on
// Taylor-Green 128^3 runs in 1.28 ms (FP64) / 1.24 ms (mixed) per step instead of 1.43 / 1.40 ms
// with the body written in the kernel itself (same arithmetic, different register allocation).
template <int D, bool DENS, bool DT, typename R, int G, bool FILT = false>
__global__ void __launch_bounds__(kBlock) k_density_dt(Fields f, Geo g, Phys ph, DevTable tab, int n, Scalars* sc,
                                                       int* overflow, int ldsTable) {
  with_table<R>(tab, DENS ? ldsTable : 0, [&](auto coef) {
    density_dt_node<D, DENS, DT, R, G, FILT>(f, g, ph, tab, coef, n, sc, overflow);
  });
}

// Integrator::selectDt (Integrator/Integrator.cc:116-169), one thread per run.
__global__ void k_select_dt(Scalars* sc, int nruns, Phys ph, double dtMin, double dtMax, double dtGrowth, int useGravity) {
  const int r = blockIdx.x*blockDim.x + threadIdx.x;
  if (r >= nruns) return;
  Scalars& s = sc[r];
  s.skinD = 0ull;
  s.skinG = 0ull;
  if (!s.active) { s.dt = 0.0; s.hdt = 0.0; return; }
  const double vote = ord_value(s.dtBits);
  const double hydro = vote*ph.cfl;
  s.hydroDt = hydro;
  const double lo = fmin(dtMin, s.tEnd - s.t), hi = fmin(dtMax, s.tEnd - s.t);
  double dt = hi;
  if (hydro > 0.0 && hydro < dt) dt = hydro;
  if (useGravity) {
    const double gd = s.gravDt;
    if (gd > 0.0 && gd < dt) dt = gd;
  }
  dt = fmin(dt, dtGrowth*s.lastDt);
  dt = fmin(hi, fmax(lo, dt));
  s.dt = dt;
  s.hdt = 0.5*dt;
}

__global__ void k_reset_dt(Scalars* sc, int nruns) {
  const int r = blockIdx.x*blockDim.x + threadIdx.x;
  if (r < nruns) sc[r].dtBits = 0xfff0000000000000ull;   // ord_bits(+inf)
}

// Time bookkeeping after a step; a run is done when it reaches its end time.
__global__ void k_advance_time(Scalars* sc, int nruns) {
  const int r = blockIdx.x*blockDim.x + threadIdx.x;
  if (r >= nruns) return;
  Scalars& s = sc[r];
  if (!s.active) return;
  s.t = s.t + s.dt;
  s.lastDt = s.dt;
  s.cycle = s.cycle + 1;
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
  const double hdt = sc[f.run[i]].hdt;
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
  const double cs = sqrt(fmax(0.0, ph.gamma*ph.g1*eps));
  f.rhom[i] = rho;
  f.epsm[i] = eps;
  f.hinvm[i] = h;
  f.Pm[i] = ph.g1*rho*eps;
  f.csm[i] = cs;
  double2* P = reinterpret_cast<double2*>(f.xhm) + 2*(size_t)i;
  P[0] = make_double2(f.xm[0][i], D >= 2 ? f.xm[1 % D][i] : 0.0);
  P[1] = make_double2(D >= 3 ? f.xm[2 % D][i] : 0.0, h);
}

// Per-node pressure term Prho = safeInv(omega)*P/rho^2 of the midpoint state (SPH.cc:425-426),
// written with the rest of node i's midpoint record (Fields::pkm), which also carries 1/rho
// (the viscous pressure Q/rho of every pair the node takes part in is formed with it).
template <int D>
__device__ __forceinline__ void write_mid_record(const Fields& f, int i, double om) {
  const double rho = f.rhom[i];
  const double prho = (om/(om*om + 1.0e-30))*f.Pm[i]/(rho*rho);
  f.prhom[i] = prho;
  double2* P = reinterpret_cast<double2*>(f.pkm) + 4*(size_t)i;
  P[0] = make_double2(f.vm[0][i], D >= 2 ? f.vm[1 % D][i] : 0.0);
  P[1] = make_double2(D >= 3 ? f.vm[2 % D][i] : 0.0, f.m[i]);
  P[2] = make_double2(rho, f.csm[i]);
  P[3] = make_double2(prho, 1.0/rho);
}

// The midpoint records (Fields::xhm, Fields::pkm) from the midpoint arrays.  keepOmega: the
// midpoint grad-h correction is the state's (no grad-h pass: gradhCorrection off, where it
// stays at its initial 1).
template <int D>
__global__ void __launch_bounds__(kBlock) k_pack_mid(Fields f, int n, int keepOmega) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2* P = reinterpret_cast<double2*>(f.xhm) + 2*(size_t)i;
  P[0] = make_double2(f.xm[0][i], D >= 2 ? f.xm[1 % D][i] : 0.0);
  P[1] = make_double2(D >= 3 ? f.xm[2 % D][i] : 0.0, f.hinvm[i]);
  double om = f.omegam[i];
  if (keepOmega) { om = f.omega[i]; f.omegam[i] = om; }
  write_mid_record<D>(f, i, om);
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
// Positions and h^-1 come from an (x, h^-1) record array.  MID: the midpoint pass, which
// also writes node i's midpoint record for the SPH pass (write_mid_record).  runMask
// (optional): only nodes of runs with a non-zero entry are updated.
// ---------------------------------------------------------------------------
template <int D, typename R, bool MID, int G, typename CP>
__device__ __forceinline__ void omega_node(const Fields& f, const Geo& g, const DevTable& tab, CP coef, int n,
                                           const double* xh, double* omega, const int* runMask) {
  const int i = xcd_block()*blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (runMask && !runMask[f.run[i]]) return;
  double xi[D], hid;
  load_xh<D>(xh, i, xi, hid);
  const R hi = (R)hid;
  const R hdeti = hdet_iso<D, R>(hi);
  double wsum = 0.0, gsum = 0.0;
  const int nn = f.cnt[i];
  for (int k0 = 0; k0 < nn; k0 += G) {
    int js[G], codes[G];
    ell_group<G>(f.nbr, (size_t)f.ld, k0, i, nn, js, codes);
    double xj[G][D], hjunk[G];
#pragma unroll
    for (int u = 0; u < G; ++u) load_xh<D>(xh, js[u], xj[u], hjunk[u]);
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (k0 + u < nn) {
        if (codes[u]) image_pos<D>(g, codes[u], xj[u]);
        R s = (R)0;
#pragma unroll
        for (int a = 0; a < D; ++a) { const R rr = hi*(R)(xi[a] - xj[u][a]); s += rr*rr; }
        const R eta = sqrt(s);
        R w, gw, graw;
        kernel_wgw_c<R>(coef, tab, eta, hdeti, w, gw, graw);
        wsum += (double)w;
        gsum += (double)(eta*gw);
      }
    }
  }
  double om;
  if (nn == 0) {
    om = 1.0;
  } else {
    wsum += (double)hdeti*tab.W0;
    om = fmax(1.0e-30, -gsum/(double(D)*wsum));
  }
  omega[i] = om;
  if (MID) write_mid_record<D>(f, i, om);
}
template <int D, typename R, bool MID, int G = kGroup>
__global__ void __launch_bounds__(kBlock) k_omega(Fields f, Geo g, DevTable tab, int n, const double* xh,
                                                  double* omega, const int* runMask, int ldsTable) {
  with_table<R>(tab, ldsTable, [&](auto coef) { omega_node<D, R, MID, G>(f, g, tab, coef, n, xh, omega, runMask); });
}

// SynchronousRK2 with TreeGravity as the only package (Integrator/SynchronousRK2.cc:40-135): the
// half step with the start-of-step derivatives (DxDt = v, DvDt = the tree acceleration), then the
// full step from the start state with the midpoint ones (TreeGravity.cc:303-316 sets DxDt = v).
template <int D>
__global__ void __launch_bounds__(kBlock) k_grav_mid(Fields f, const Scalars* sc, int n) {
  const int i = xcd_block()*blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double hdt = sc[f.run[i]].hdt;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const double v = f.v[a][i];
    f.xm[a][i] = f.x[a][i] + hdt*v;
    f.vm[a][i] = v + hdt*f.DvDtG[a][i];
  }
}
template <int D>
__global__ void __launch_bounds__(kBlock) k_grav_full(Fields f, const Scalars* sc, int n) {
  const int i = xcd_block()*blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dt = sc[f.run[i]].dt;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const double vm = f.vm[a][i];
    f.x[a][i] = f.x[a][i] + dt*vm;
    f.v[a][i] = f.v[a][i] + dt*f.DvDtG[a][i];
    f.DxDt[a][i] = vm;
  }
}

// One pass of iterateIdealH (Utilities/iterateIdealH.cc:118-176) with the SPHSmoothingScale
// package alone (SmoothingScale/SPHSmoothingScale.cc:154-245) on the connectivity state: the
// zeroth moment of kernelValueSPH over the node's neighbours (images included) gives the ideal h;
// a node not yet converged takes it, and is flagged done once its H moved by at most tol.  The
// flags are indexed by the node's run-local id, so they survive the reorders between passes.
template <int D>
__global__ void __launch_bounds__(kBlock) k_ideal_h(Fields f, Geo g, Phys ph, DevTable tab, int n, double tol,
                                                    int* done, unsigned long long* maxDelta) {
  const int i = xcd_block()*blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t id = f.oid[i] & (((int64_t)1 << 40) - 1);
  if (done[id]) return;
  const double* coef = table_coef<double>(tab);
  double xi[D], hi;
  load_xh<D>(f.xh, i, xi, hi);
  const size_t ld = (size_t)f.ld;
  double m0 = 0.0;
  const int nn = f.cnt[i];
  for (int k = 0; k < nn; ++k) {
    const uint32_t e = f.nbr[ell_at(k, i, ld)];
    const int j = (int)(e & kIdxMask), code = (int)(e >> kImgShift);
    double xj[D], hj;
    load_xh<D>(f.xh, j, xj, hj);
    if (code) image_pos<D>(g, code, xj);
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) { const double r = hi*(xi[a] - xj[a]); s += r*r; }
    double w, gw, graw;
    kernel_wgw_c<double>(coef, tab, sqrt(s), hdet_iso<D, double>(hi), w, gw, graw);
    m0 += fabs(graw);
  }
  m0 = rootnu<D>(fmax(0.0, m0));
  const double cur = (fabs(m0) <= 1.0e-15*fmax(1.0, fabs(m0))) ? 0.5*ph.nperh : nperh_lookup(tab, m0);
  const double sr = fmin(4.0, fmax(0.25, ph.nperh/(cur + 1.0e-30)));
  const double aa = sr < 1.0 ? 0.4*(1.0 + sr*sr) : 0.4*(1.0 + 1.0/(sr*sr*sr));
  const double hi1 = fmin(ph.hmax, fmax(ph.hmin, (1.0/hi)*(1.0 - aa + aa*sr)));
  const double H1 = 1.0/hi1;
  // eigenvalues of H1^1/2 H^-1 H1^1/2 (isotropic: one value)
  const double sq = sqrt(H1);
  const double dH = fabs((sq*(1.0/hi))*sq - 1.0);
  if (dH <= tol) done[id] = 1;
  const unsigned long long b = (unsigned long long)__double_as_longlong(dH);
  if (b > *maxDelta) atomicMax(maxDelta, b);
  f.hinv[i] = H1;
}

// Per-run extents for the next grid: min and max hinv and the bounding box, as ordered bits.
// One block per tile (run, begin, end) of a run's contiguous particle segment; a block
// reduces its tile and adds one set of atomics.
// ext layout per run (kExt values): [hinv_min, xmin0, xmax0, xmin1, xmax1, xmin2, xmax2, hinv_max].
constexpr int kExt = 8;
__host__ __device__ __forceinline__ bool ext_is_min(int q) { return q == 0 || (q < 7 && (q & 1)); }
template <int D>
__global__ void __launch_bounds__(kBlock) k_extent(Fields f, const int* tiles, unsigned long long* ext) {
  __shared__ unsigned long long red[kExt][kBlock/64];
  const int r = tiles[3*blockIdx.x], b = tiles[3*blockIdx.x + 1], e = tiles[3*blockIdx.x + 2];
  unsigned long long v[kExt];
#pragma unroll
  for (int q = 0; q < kExt; ++q) v[q] = ext_is_min(q) ? ~0ull : 0ull;
  for (int i = b + threadIdx.x; i < e; i += blockDim.x) {
    const unsigned long long h = ord_bits(f.hinv[i]);
    v[0] = h < v[0] ? h : v[0];
    v[7] = h > v[7] ? h : v[7];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const unsigned long long x = ord_bits(f.x[a][i]);
      v[1 + 2*a] = x < v[1 + 2*a] ? x : v[1 + 2*a];
      v[2 + 2*a] = x > v[2 + 2*a] ? x : v[2 + 2*a];
    }
  }
#pragma unroll
  for (int q = 0; q < kExt; ++q) {
    unsigned long long x = v[q];
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long y = __shfl_xor(x, o);
      x = ext_is_min(q) ? (y < x ? y : x) : (y > x ? y : x);
    }
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long* out = ext + (size_t)r*kExt;
    for (int q = 0; q < kExt; ++q) {
      if (q >= 1 + 2*D && q < 7) continue;
      unsigned long long x = red[q][0];
      for (int w = 1; w < kBlock/64; ++w) {
        const unsigned long long y = red[q][w];
        x = ext_is_min(q) ? (y < x ? y : x) : (y > x ? y : x);
      }
      if (ext_is_min(q)) atomicMin(&out[q], x);
      else atomicMax(&out[q], x);
    }
  }
}

__global__ void k_reset_extent(unsigned long long* ext, int nruns) {
  const int k = blockIdx.x*blockDim.x + threadIdx.x;
  if (k >= nruns*kExt) return;
  ext[k] = ext_is_min(k % kExt) ? ~0ull : 0ull;
}

// Non-finite guard on the state (NaN/Inf -> per-run status word).
template <int D>
__global__ void __launch_bounds__(kBlock) k_finite(Fields f, int n, Scalars* sc) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool ok = isfinite(f.rho[i]) && isfinite(f.eps[i]) && isfinite(f.hinv[i]);
  for (int a = 0; a < D; ++a) ok = ok && isfinite(f.x[a][i]) && isfinite(f.v[a][i]);
  if (!ok) atomicOr(&sc[f.run[i]].nonfinite, 1);
}

// dst[0..n) = value.  Clears flags inside captured step graphs, which hold no memset nodes.
__global__ void k_fill_int(int* dst, int n, int value) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i < n) dst[i] = value;
}

// dst[0..n) = src[0..n) (doubles): the velocities of the finished runs behind the active ones,
// which the update does not write, carried into the other velocity buffer.
__global__ void __launch_bounds__(kBlock) k_copy_doubles(const double* __restrict__ src, double* __restrict__ dst,
                                                         int64_t n) {
  const int64_t i = (int64_t)blockIdx.x*blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Unique-pair count of the neighbour lists (ConnectivityMap::nodePairList().size(),
// Neighbor/ConnectivityMap.cc:925-927): an entry (i, j) is i's pair when j is an image or a later
// slot (calculatePairInteraction, ConnectivityMapInline.hh:305-307).  One partial count per block
// (no atomics), summed by k_sum_partials.
__global__ void __launch_bounds__(kBlock) k_count_pairs(const uint32_t* nbr, const int32_t* cnt, size_t ld, int n,
                                                        unsigned long long* partial) {
  __shared__ unsigned long long red[kBlock/64];
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  unsigned long long c = 0;
  if (i < n) {
    const int nn = cnt[i];
    for (int k = 0; k < nn; ++k) {
      const uint32_t e = nbr[ell_at(k, i, ld)];
      c += ((e >> kImgShift) != 0u || (int)(e & kIdxMask) > i) ? 1ull : 0ull;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    partial[blockIdx.x] = t;
  }
}

// out[0] = sum of partial[0..nb); one block.
__global__ void __launch_bounds__(kBlock) k_sum_partials(const unsigned long long* partial, int nb,
                                                         unsigned long long* out) {
  __shared__ unsigned long long red[kBlock/64];
  unsigned long long c = 0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) c += partial[k];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    out[0] = t;
  }
}

}  // namespace spx
