// spx_gravity.hpp -- TreeGravity on the device (3-D octree; 2-D quadtree).
//
// Tree: the reference's octree (Gravity/TreeGravityInline.hh:37-160, TreeGravity.cc:362-477)
// rebuilt from 60-bit Morton keys.  The level-l cell of a point is the l-level prefix of its
// key (floor(u*2^l) with u = (x - xmin)/boxLength is the prefix of floor(u*2^20), since the
// scaling is by a power of two), and the reference's split rule -- a cell is subdivided when a
// second point arrives, down to level 20 where points are listed -- gives: cells at level l are
// the l-prefix groups of points whose (l-1)-group holds at least two points.  Point p owns the
// cells (p, l) for l from 1 + lcp(p-1, p) to its leaf level min(20, 1 + max lcp with its
// sorted neighbours); numbering them by (p, l) is a pre-order walk, so each cell's subtree is
// the contiguous index range [c, skip[c]) and the force walk needs no stack.
//
// Walk: TreeGravity::applyTreeForces (TreeGravity.cc:818-950) per point from the root's
// daughters: accept a cell's monopole when |xcm - xi|^2 > cellsize^2/opening^2 + rcm2cc2,
// otherwise sum a leaf's points directly (self screened by r^2/soft^2 > 1e-10) or descend.
// Force M/(r^2+soft^2) along the unit vector, potential -M/sqrt(r^2+soft^2) in 3-D; in 2-D
// M/sqrt(r^2+soft^2) and -M log(sqrt(r^2+soft^2)) (TreeGravity.cc:54-90); then DvDt *= G,
// potential *= G, dt vote ftimestep*min sqrt(soft/|a|)/sqrt(G) (TreeGravity.cc:143-328,483-497).
//
// 2-D: the points get z = 0, so the interleaved keys' z bits are all zero and the octree's cells
// are exactly the quadtree's (level-l cell = floor(u*2^l) in x and y, TreeGravityInline.hh:16-33);
// the cell-centre offset rcm2cc2 is summed over x and y only.
#pragma once

#include <rocprim/rocprim.hpp>

#include <vector>

#include "spx_device.hpp"
#include "spx_error.hpp"

namespace spx {

constexpr int kTreeBits = 20;   // deepest level (num1dbits - 1)

struct TreeView {
  const double* tx[3];     // tree-order positions
  const double* tm;        // tree-order masses
  const uint64_t* code;    // sorted keys
  const int* first;        // cell point range [first, last]
  const int* last;
  const int* skip;
  const unsigned char* level;
  const double* M;
  const double* cx[3];     // centre of mass
  const double* rc2;       // |xcm - cell centre|^2
  int ncells, n;
  double xmin[3], box;
};

__device__ __forceinline__ int lcp_levels(uint64_t a, uint64_t b) {
  const uint64_t x = a ^ b;
  if (x == 0) return kTreeBits;
  const int lz = __clzll(x) - (64 - 3*kTreeBits);   // leading equal bits within the 60-bit key
  return lz/3;
}

__global__ void __launch_bounds__(kBlock) k_tree_bbox(double* const* xs, int n, double* part) {
  __shared__ double red[kBlock];
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  for (int q = 0; q < 6; ++q) {
    const int a = q >> 1;
    double v = 1.7976931348623157e308;
    if (i < n) v = (q & 1) ? -xs[a][i] : xs[a][i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock/2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x*6 + q] = red[0];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kBlock) k_tree_codes(double* const* xs, int n, double x0, double y0, double z0, double box,
                                                       uint64_t* codes, uint32_t* idx) {
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double o[3] = {x0, y0, z0};
  const uint64_t ncell = 1ull << kTreeBits;
  uint64_t key = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    uint64_t c = 0;
    if (box > 0.0) {
      const double u = (xs[a][i] - o[a])/box*(double)ncell;      // TreeGravityInline.hh buildCellKey
      c = u <= 0.0 ? 0 : (uint64_t)u;
      c = c < ncell - 1 ? c : ncell - 1;
    }
    key |= spread_bits(c, 3) << a;
  }
  codes[i] = key;
  idx[i] = (uint32_t)i;
}

// Leaf level and first owned level of each sorted point -> number of cells it heads.
__global__ void __launch_bounds__(kBlock) k_tree_heads(const uint64_t* code, int n, int* nheads, unsigned char* lstart) {
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int lp = p > 0 ? lcp_levels(code[p - 1], code[p]) : -1;
  const int ln = p + 1 < n ? lcp_levels(code[p], code[p + 1]) : -1;
  int leaf = 1 + (lp > ln ? lp : ln);
  leaf = leaf < kTreeBits ? leaf : kTreeBits;
  if (n == 1) leaf = 0;
  const int ls = p == 0 ? 0 : 1 + lp;
  nheads[p] = leaf >= ls ? leaf - ls + 1 : 0;
  lstart[p] = (unsigned char)(ls > 63 ? 63 : ls);
}

// Emit the cells (p, l): level, point range, skip pointer.
__global__ void __launch_bounds__(kBlock) k_tree_cells(const uint64_t* code, int n, const int* cstart, const int* nheads,
                                                       const unsigned char* lstart, int ncells, int* first, int* last,
                                                       int* skip, unsigned char* level) {
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int k0 = cstart[p], nh = nheads[p], ls = lstart[p];
  for (int k = 0; k < nh; ++k) {
    const int l = ls + k;
    const int c = k0 + k;
    // last point sharing the l-prefix: binary search over the sorted keys
    const int sh = 3*(kTreeBits - l);
    const uint64_t pre = l == 0 ? 0 : (code[p] >> sh);
    int lo = p, hi = n - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      const uint64_t pm = l == 0 ? 0 : (code[mid] >> sh);
      if (pm == pre) lo = mid; else hi = mid - 1;
    }
    first[c] = p;
    last[c] = lo;
    level[c] = (unsigned char)l;
    skip[c] = lo + 1 < n ? cstart[lo + 1] : ncells;
  }
}

// Monopoles of the cells at one level, bottom-up (children already done).
__global__ void __launch_bounds__(kBlock) k_tree_moments(TreeView t, int lvl, double* M, double* cx0, double* cx1, double* cx2,
                                                         double* rc2, int dim) {
  const int c = blockIdx.x*blockDim.x + threadIdx.x;
  if (c >= t.ncells || t.level[c] != lvl) return;
  const int f = t.first[c], l = t.last[c];
  double m = 0.0, sx = 0.0, sy = 0.0, sz = 0.0, xc[3];
  if (f == l) {
    m = t.tm[f];
    xc[0] = t.tx[0][f]; xc[1] = t.tx[1][f]; xc[2] = t.tx[2][f];
  } else {
    if (lvl == kTreeBits) {
      for (int q = f; q <= l; ++q) {
        const double mq = t.tm[q];
        m += mq; sx += mq*t.tx[0][q]; sy += mq*t.tx[1][q]; sz += mq*t.tx[2][q];
      }
    } else {
      const int end = t.skip[c];
      for (int d = c + 1; d < end; d = t.skip[d]) {
        const double md = M[d];
        m += md; sx += md*cx0[d]; sy += md*cx1[d]; sz += md*cx2[d];
      }
    }
    xc[0] = sx/m; xc[1] = sy/m; xc[2] = sz/m;
  }
  M[c] = m;
  cx0[c] = xc[0]; cx1[c] = xc[1]; cx2[c] = xc[2];
  const double cs = t.box/(double)(1u << lvl);
  const uint64_t key = t.code[f];
  double r2 = 0.0;
  for (int a = 0; a < dim; ++a) {
    uint64_t ia = 0;
    for (int b = 0; b < lvl; ++b) ia |= ((key >> (3*(kTreeBits - lvl + b) + a)) & 1ull) << b;
    const double cc = t.xmin[a] + ((double)ia + 0.5)*cs;
    const double d = xc[a] - cc;
    r2 += d*d;
  }
  rc2[c] = r2;
}

__global__ void __launch_bounds__(kBlock) k_tree_gather(double* const* xs, const double* m, const uint32_t* perm, int n,
                                                        double* tx0, double* tx1, double* tx2, double* tm, double2* tp) {
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t q = perm[p];
  const double x = xs[0][q], y = xs[1][q], z = xs[2][q], mq = m[q];
  tx0[p] = x; tx1[p] = y; tx2[p] = z;
  tm[p] = mq;
  // the walk's leaf record: (x, y), (z, m) in one 32-byte block
  tp[2*(size_t)p] = make_double2(x, y);
  tp[2*(size_t)p + 1] = make_double2(z, mq);
}

// A cell as the walk reads it: one 64-byte record, (x, y), (z, M) of the centre of mass, the
// opening radius^2 l^2/theta^2 + |xcm - centre|^2 for the walk's opening angle, and (first,
// last, skip, level) -- one cache line per visited cell instead of one per array.
struct alignas(64) TreeCell {
  double x, y, z, M;
  double open, pad;
  int first, last, skip, level;
};
static_assert(sizeof(TreeCell) == 64, "TreeCell is one 64-byte record");

__global__ void __launch_bounds__(kBlock) k_tree_pack(TreeView v, const double* cs2, double open2, TreeCell* cells) {
  const int c = blockIdx.x*blockDim.x + threadIdx.x;
  if (c >= v.ncells) return;
  const int lvl = v.level[c];
  TreeCell t;
  t.x = v.cx[0][c];
  t.y = v.cx[1][c];
  t.z = v.cx[2][c];
  t.M = v.M[c];
  t.open = cs2[lvl]/open2 + v.rc2[c];
  t.pad = 0.0;
  t.first = v.first[c];
  t.last = v.last[c];
  t.skip = v.skip[c];
  t.level = lvl;
  cells[c] = t;
}

// Lanes that walk the tree together (64: the wavefront; 16, 8, 4: aligned lane groups).
// Measured on the Evrard collapse (250k points) with the cells read as separate arrays: 64 ->
// 5.0 ms per walk, 16 -> 4.15, 8 -> 3.99, 4 -> 4.69; with the packed cell records (gravity per
// step over 60 steps from the start): 16 -> 2.16 ms, 8 -> 1.99, 4 -> 1.86, 2 -> 1.92, 1 -> 2.14;
// with single-point leaves taken through the monopole path (one branch, no point-record read):
// 8 -> 1.72 ms, 4 -> 1.59, 2 -> 1.60.  Same bits for every group size.
constexpr int kTreeGroup = 4;

// Device-library wavefront minimum (DPP lane exchanges, no LDS round trip).
extern "C" __device__ int __ockl_wfred_min_i32(int);

// Minimum over each aligned group of GROUP (4, 8 or 16) lanes by DPP lane exchanges; every
// lane of the group ends with the group's minimum.  Quad permutes [1,0,3,2] and [2,3,0,1]
// combine the lanes of a quad, the half-row mirror (lane l <-> 7 - l within 8) pairs the two
// quads of a half-row, the row mirror (l <-> 15 - l within 16) the two halves of a row.
template <int GROUP>
__device__ __forceinline__ int group_min(int v) {
  v = min(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xf, 0xf, false));
  if (GROUP >= 8) v = min(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xf, 0xf, false));
  if (GROUP >= 16) v = min(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xf, 0xf, false));
  return v;
}

// applyTreeForces for the point in tree slot p; results scattered to its hydro slot.
// The wavefront walks the pre-order cells together: each lane keeps the next cell its own
// walk would visit, the wavefront visits the smallest of them, and only the lanes waiting
// on that cell act on it.  Every lane therefore takes exactly the decisions, in exactly the
// order, of an independent stackless walk (same bits).  The visited cell is wavefront-
// uniform (readfirstlane), so its data and its leaf points are read through the scalar
// cache; the tree arrays come as separate restrict-qualified arguments, which tells the
// compiler they are not written here.  GROUP 16: each row of 16 lanes walks together instead
// of the whole wavefront (the walks of 16 neighbouring points differ less than those of 64, so
// fewer cells are visited), with per-group cell reads (64/GROUP addresses per instruction);
// 8 and 4 likewise.  COUNT (a measurement variant, never in a step): also counts each point's
// monopole (accepted cell) and point-point interactions into counts[2p], counts[2p+1].
template <int GROUP, bool COUNT = false, int DIM = 3>
__global__ void __launch_bounds__(kBlock) k_tree_walk(
    const double* __restrict__ tx0, const double* __restrict__ tx1, const double* __restrict__ tx2,
    const double2* __restrict__ tp, const TreeCell* __restrict__ cells, int ncells, int npts,
    const uint32_t* __restrict__ perm, double soft2, double soft, double G, double* __restrict__ acc0,
    double* __restrict__ acc1, double* __restrict__ acc2, double* __restrict__ pot, double* __restrict__ blockMin,
    unsigned long long* __restrict__ counts = nullptr) {
  const int p = blockIdx.x*blockDim.x + threadIdx.x;
  unsigned long long nCell = 0, nPoint = 0;
  const bool valid = p < npts;
  double dtv = 1.7976931348623157e308;
  double xi = 0.0, yi = 0.0, zi = 0.0;
  if (valid) { xi = tx0[p]; yi = tx1[p]; zi = tx2[p]; }
  double ax = 0.0, ay = 0.0, az = 0.0, phi = 0.0;
  int next = valid ? 1 : ncells;
  // every iteration moves the group's smallest pending cell strictly forward, so a walk
  // ends within ncells iterations; the bound only guards the loop
  for (int guard = 0; guard <= ncells; ++guard) {
    int c;
    if (GROUP == 64) {
      c = __builtin_amdgcn_readfirstlane(__ockl_wfred_min_i32(next));
      if (c >= ncells) break;                     // every lane of the wavefront is done
    } else {
      c = group_min<GROUP>(next);
      if (__all(c >= ncells)) break;              // every group is done
      if (c >= ncells) continue;                  // this group is done
    }
    // the cell's record, read at one address per group
    const double2* cr = reinterpret_cast<const double2*>(cells + c);
    const double2 c01 = cr[0], c23 = cr[1], co = cr[2];
    const int4 ci = *reinterpret_cast<const int4*>(cr + 3);
    const double ccx = c01.x, ccy = c01.y, ccz = c23.x, Mc = c23.y, open = co.x;
    const int q0 = ci.x, q1 = ci.y, sk = ci.z, lvl = ci.w;
    if (next != c) continue;
    const double dx = ccx - xi, dy = ccy - yi, dz = ccz - zi;
    const double r2 = dx*dx + dy*dy + dz*dz;
    // An accepted cell acts through its monopole.  A single-point leaf acts as its point, and
    // its record already holds that point: the monopole pass copies a one-point cell's x and m
    // unchanged (no m*x/m round trip), so (ccx, ccy, ccz, Mc) are the point's own values and
    // the point-point term below is the monopole formula with the point's softening guard.
    // One code path for both: no second, dependent read of the point record, and lanes of a
    // wavefront at accepted cells and at single-point leaves no longer run two branches (gravity
    // 1.86 -> 1.59 ms per step on the 250k Evrard collapse; results move at roundoff, 7e-15 in
    // density after 65 steps, as the compiler contracts the merged arithmetic differently).
    const bool mono = r2 > open;
    const bool single = !mono && q0 == q1;
    if (mono || (single && r2/soft2 > 1.0e-10)) {
      // 1/|r| and 1/sqrt(r^2 + s^2) from reciprocal square roots (the force is M/(r^2 + s^2) along
      // r/|r|, the potential -M/sqrt(r^2 + s^2)): 11 % faster than two square roots and three
      // divisions on the 250k Evrard walk, results within 2e-16 (L1 of density after 65 steps)
      const bool far = single || r2 > 1.0e-100;
      const double inv = far ? rsqrt(r2) : 0.0;
      const double nx = far ? dx*inv : 1.0, ny = dy*inv, nz = dz*inv;
      const double rs = r2 + soft2;
      const double ri = rsqrt(rs);
      const double f = DIM == 3 ? Mc*(ri*ri) : Mc*ri;
      ax += f*nx; ay += f*ny; az += f*nz;
      phi -= DIM == 3 ? Mc*ri : Mc*(0.5*log(rs));
      if (COUNT) { if (mono) ++nCell; else ++nPoint; }
    }
    if (mono || single) {
      next = sk;
    } else if (lvl == kTreeBits) {
      for (int q = q0; q <= q1; ++q) {
        const double2 pa = tp[2*(size_t)q], pb = tp[2*(size_t)q + 1];
        const double ex = pa.x - xi, ey = pa.y - yi, ez = pb.x - zi;
        const double q2 = ex*ex + ey*ey + ez*ez;
        if (q2/soft2 > 1.0e-10) {
          const double inv = rsqrt(q2);
          const double rs = q2 + soft2;
          const double mq = pb.y;
          const double ri = rsqrt(rs);
          const double f = DIM == 3 ? mq*(ri*ri) : mq*ri;
          ax += f*(ex*inv); ay += f*(ey*inv); az += f*(ez*inv);
          phi -= DIM == 3 ? mq*ri : mq*(0.5*log(rs));
          if (COUNT) ++nPoint;
        }
      }
      next = sk;
    } else {
      next = c + 1;
    }
  }
  if (valid) {
    const double amag = sqrt(ax*ax + ay*ay + az*az);
    if (amag > 0.0) dtv = sqrt(soft/amag);
    const uint32_t s = perm[p];
    acc0[s] = G*ax;
    acc1[s] = G*ay;
    acc2[s] = G*az;
    pot[s] = G*phi;
    if (COUNT) { counts[2*(size_t)p] = nCell; counts[2*(size_t)p + 1] = nPoint; }
  }
  __shared__ double red[kBlock];
  red[threadIdx.x] = dtv;
  __syncthreads();
  for (int s = kBlock/2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) blockMin[blockIdx.x] = red[0];
}

__global__ void k_tree_dt(const double* blockMin, int nb, Scalars* sc, double G, double ftimestep) {
  __shared__ double red[kBlock];
  double m = 1.7976931348623157e308;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) m = fmin(m, blockMin[b]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kBlock/2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sc->dtMinAcc = red[0]/sqrt(G);
    sc->gravDt = ftimestep*sc->dtMinAcc;
  }
}

class GravityTree {
 public:
  int D = 3;
  int64_t n = 0;
  hipStream_t stream = nullptr;
  bool active = false;
  std::vector<void*> allocs;
  int64_t bytes = 0;                // device memory the tree holds
  uint64_t *codes = nullptr, *codesSorted = nullptr;
  uint32_t *idx = nullptr, *perm = nullptr;
  double *tx[3] = {nullptr, nullptr, nullptr}, *tm = nullptr;
  double2* tp = nullptr;            // tree-order (x, y), (z, m) records of the walk's leaves
  TreeCell* cells = nullptr;        // the walk's cell records
  double packedOpen2 = -1.0;        // opening^2 the cell records were packed for (< 0: stale)
  int *nheads = nullptr, *cstart = nullptr;
  unsigned char* lstart = nullptr;
  void* tmp = nullptr;
  size_t tmpBytes = 0;
  double* part = nullptr;
  double* hpart = nullptr;
  double* cs2 = nullptr;
  double** xsPtrs = nullptr;
  double* zero = nullptr;           // 2-D: the z coordinate of every point
  double* sink = nullptr;           // 2-D: the walk's z acceleration
  // cells
  int64_t cellCap = 0;
  int ncells = 0;
  int *first = nullptr, *last = nullptr, *skip = nullptr;
  unsigned char* level = nullptr;
  double *M = nullptr, *cx[3] = {nullptr, nullptr, nullptr}, *rc2 = nullptr;
  TreeView view{};

  template <typename T>
  T* alloc(size_t count) {
    void* p = nullptr;
    const size_t b = std::max<size_t>(count, 1)*sizeof(T);
    HIP_CHECK(hipMalloc(&p, b));
    allocs.push_back(p);
    bytes += (int64_t)b;
    return static_cast<T*>(p);
  }

  void init(int dim, int64_t n_, int ld, hipStream_t s, bool enabled) {
    D = dim;
    n = n_;
    stream = s;
    active = enabled;
    if (!active) return;
    if (D != 3 && D != 2) raise(SPX_ERR_UNSUPPORTED, "TreeGravity on the GPU is implemented for 2-D and 3-D");
    (void)ld;
    if (D == 2) {
      // z = 0 for every point (the quadtree as a flat octree), and a sink for the z acceleration
      zero = alloc<double>(n);
      HIP_CHECK(hipMemsetAsync(zero, 0, (size_t)std::max<int64_t>(n, 1)*sizeof(double), stream));
      sink = alloc<double>(n);
    }
    codes = alloc<uint64_t>(n);
    codesSorted = alloc<uint64_t>(n);
    idx = alloc<uint32_t>(n);
    perm = alloc<uint32_t>(n);
    for (int a = 0; a < 3; ++a) tx[a] = alloc<double>(n);
    tm = alloc<double>(n);
    tp = alloc<double2>(2*(size_t)n);
    nheads = alloc<int>(n + 1);
    cstart = alloc<int>(n + 1);
    lstart = alloc<unsigned char>(n);
    const int nb = (int)((n + kBlock - 1)/kBlock);
    part = alloc<double>((size_t)nb*6 + 8);
    HIP_CHECK(hipHostMalloc((void**)&hpart, ((size_t)nb*6 + 8)*sizeof(double), hipHostMallocDefault));
    cs2 = alloc<double>(32);
    xsPtrs = alloc<double*>(3);
    size_t t1 = 0, t2 = 0;
    HIP_CHECK(rocprim::radix_sort_pairs(nullptr, t1, codes, codesSorted, idx, perm, (size_t)n, 0, 60, stream));
    HIP_CHECK(rocprim::exclusive_scan(nullptr, t2, nheads, cstart, 0, (size_t)n + 1, rocprim::plus<int>(), stream));
    tmpBytes = std::max(t1, t2);
    tmp = alloc<char>(tmpBytes);
  }

  void release() {
    for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
    if (hpart) (void)hipHostFree(hpart);
    hpart = nullptr;
  }

  void ensureCells(int64_t need) {
    if (need <= cellCap) return;
    const int64_t cap = need + need/4 + 1024;
    first = alloc<int>(cap);
    last = alloc<int>(cap);
    skip = alloc<int>(cap);
    level = alloc<unsigned char>(cap);
    M = alloc<double>(cap);
    for (int a = 0; a < 3; ++a) cx[a] = alloc<double>(cap);
    rc2 = alloc<double>(cap);
    cells = alloc<TreeCell>(cap);
    cellCap = cap;
  }

  void build(double* const* xs, const double* m, int nn) {
    const int nb = (int)((nn + kBlock - 1)/kBlock);
    double* hx[3] = {xs[0], xs[1], D == 3 ? xs[2] : zero};
    HIP_CHECK(hipMemcpyAsync(xsPtrs, hx, sizeof(hx), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_tree_bbox, dim3(nb), dim3(kBlock), 0, stream, xsPtrs, nn, part);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(hpart, part, (size_t)nb*6*sizeof(double), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) { lo[a] = 1.7976931348623157e308; hi[a] = -1.7976931348623157e308; }
    for (int b = 0; b < nb; ++b)
      for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], hpart[b*6 + 2*a]); hi[a] = std::max(hi[a], -hpart[b*6 + 2*a + 1]); }
    double box = 0.0;
    for (int a = 0; a < 3; ++a) box = std::max(box, std::abs(hi[a] - lo[a]));
    hipLaunchKernelGGL(k_tree_codes, dim3(nb), dim3(kBlock), 0, stream, xsPtrs, nn, lo[0], lo[1], lo[2], box, codes, idx);
    HIP_CHECK(hipGetLastError());
    size_t t = tmpBytes;
    HIP_CHECK(rocprim::radix_sort_pairs(tmp, t, codes, codesSorted, idx, perm, (size_t)nn, 0, 60, stream));
    hipLaunchKernelGGL(k_tree_gather, dim3(nb), dim3(kBlock), 0, stream, xsPtrs, m, perm, nn, tx[0], tx[1], tx[2], tm,
                       tp);
    packedOpen2 = -1.0;
    hipLaunchKernelGGL(k_tree_heads, dim3(nb), dim3(kBlock), 0, stream, codesSorted, nn, nheads, lstart);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemsetAsync(nheads + nn, 0, sizeof(int), stream));
    t = tmpBytes;
    HIP_CHECK(rocprim::exclusive_scan(tmp, t, nheads, cstart, 0, (size_t)nn + 1, rocprim::plus<int>(), stream));
    int total = 0;
    HIP_CHECK(hipMemcpyAsync(&total, cstart + nn, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    ncells = total;
    ensureCells(total);
    hipLaunchKernelGGL(k_tree_cells, dim3(nb), dim3(kBlock), 0, stream, codesSorted, nn, cstart, nheads, lstart, total,
                       first, last, skip, level);
    HIP_CHECK(hipGetLastError());
    view.tx[0] = tx[0]; view.tx[1] = tx[1]; view.tx[2] = tx[2];
    view.tm = tm;
    view.code = codesSorted;
    view.first = first; view.last = last; view.skip = skip; view.level = level;
    view.M = M; view.cx[0] = cx[0]; view.cx[1] = cx[1]; view.cx[2] = cx[2]; view.rc2 = rc2;
    view.ncells = total;
    view.n = nn;
    for (int a = 0; a < 3; ++a) view.xmin[a] = lo[a];
    view.box = box;
    const int cb = (int)((total + kBlock - 1)/kBlock);
    for (int l = kTreeBits; l >= 0; --l) {
      hipLaunchKernelGGL(k_tree_moments, dim3(cb), dim3(kBlock), 0, stream, view, l, M, cx[0], cx[1], cx[2], rc2, D);
    }
    HIP_CHECK(hipGetLastError());
    double h2[32];
    for (int l = 0; l < 32; ++l) h2[l] = l <= kTreeBits ? (box*box)/(double)(1ull << (2*l)) : 0.0;
    HIP_CHECK(hipMemcpyAsync(cs2, h2, sizeof(h2), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  }

  // The walk's cell records for this opening angle (once per tree and angle).
  void pack(double opening) {
    const double open2 = opening*opening;
    if (packedOpen2 == open2) return;
    const int cb = (int)((view.ncells + kBlock - 1)/kBlock);
    hipLaunchKernelGGL(k_tree_pack, dim3(std::max(cb, 1)), dim3(kBlock), 0, stream, view, (const double*)cs2, open2, cells);
    HIP_CHECK(hipGetLastError());
    packedOpen2 = open2;
  }

  void evaluate(double* const* acc, double* pot, int nn, double G, double soft, double opening, Scalars* sc,
                double ftimestep) {
    const int nb = (int)((nn + kBlock - 1)/kBlock);
    double* az = D == 3 ? acc[2] : sink;
    pack(opening);
    auto walk = D == 3 ? k_tree_walk<kTreeGroup, false, 3> : k_tree_walk<kTreeGroup, false, 2>;
    hipLaunchKernelGGL(walk, dim3(nb), dim3(kBlock), 0, stream, view.tx[0], view.tx[1], view.tx[2], (const double2*)tp,
                       (const TreeCell*)cells, view.ncells, view.n, (const uint32_t*)perm, soft*soft, soft, G,
                       acc[0], acc[1], az, pot, part, (unsigned long long*)nullptr);
    hipLaunchKernelGGL(k_tree_dt, dim3(1), dim3(kBlock), 0, stream, part, nb, sc, G, ftimestep);
    HIP_CHECK(hipGetLastError());
  }

  // Interactions of one walk over the current tree (the counting variant of the walk, into
  // throw-away outputs): out[0] monopole (accepted cell) interactions, out[1] point-point
  // interactions, summed over the points.  A measurement aid: allocates, waits on the stream.
  void countInteractions(double soft, double opening, int64_t* out) {
    const int nn = (int)view.n;
    if (!active || nn <= 0) raise(SPX_ERR_STATE, "no TreeGravity evaluation to count");
    const int nb = (int)((nn + kBlock - 1)/kBlock);
    double* scratch = nullptr;
    unsigned long long* counts = nullptr;
    HIP_CHECK(hipMalloc(&scratch, (size_t)(4*(size_t)std::max(nn, 1) + nb)*sizeof(double)));
    HIP_CHECK(hipMalloc(&counts, 2*(size_t)std::max(nn, 1)*sizeof(unsigned long long)));
    double* a = scratch;
    pack(opening);
    auto walk = D == 3 ? k_tree_walk<kTreeGroup, true, 3> : k_tree_walk<kTreeGroup, true, 2>;
    hipLaunchKernelGGL(walk, dim3(nb), dim3(kBlock), 0, stream, view.tx[0], view.tx[1], view.tx[2], (const double2*)tp,
                       (const TreeCell*)cells, view.ncells, view.n, (const uint32_t*)perm, soft*soft, soft, 1.0, a, a + nn,
                       a + 2*(size_t)nn, a + 3*(size_t)nn, a + 4*(size_t)nn, counts);
    const hipError_t e = hipGetLastError();
    std::vector<unsigned long long> h(2*(size_t)nn);
    hipError_t e2 = hipSuccess;
    if (e == hipSuccess) {
      e2 = hipMemcpyAsync(h.data(), counts, h.size()*sizeof(unsigned long long), hipMemcpyDeviceToHost, stream);
      if (e2 == hipSuccess) e2 = hipStreamSynchronize(stream);
    }
    (void)hipFree(scratch);
    (void)hipFree(counts);
    HIP_CHECK(e);
    HIP_CHECK(e2);
    out[0] = out[1] = 0;
    for (int p = 0; p < nn; ++p) { out[0] += (int64_t)h[2*(size_t)p]; out[1] += (int64_t)h[2*(size_t)p + 1]; }
  }
};

}  // namespace spx
