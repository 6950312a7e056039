// Density-weighted K-means interpolation-point selection for gfx950.
//
// Reproduces KMEAN of LibISDF (src/common/utility.cpp:631-988) at one rank, bit
// for bit: every FP64 expression keeps the reference's evaluation order and
// rounding (no FMA contraction: this file is built with -ffp-contract=off and
// the pragma below), and every order-dependent sum runs in the reference's
// order.  The parallel structure is the GPU's own:
//
//  * weights: one lane per grid point, states streamed in ascending order
//    (coalesced along the column-major grid; HBM-bound);
//  * assignment: distance + first-minimum argmin fused, the distance matrix is
//    never stored; centroids are wave-uniform so they come through the scalar
//    cache;
//  * centroid update: a stable counting sort of point indices by label (chunk
//    histograms, scans, wave-ballot ranking), then one wave per cluster sums
//    its members in ascending point order: the same sequence of roundings as
//    the reference's single loop over points (utility.cpp:816-822);
//  * the loop runs on the device behind an "active" flag, so the host only
//    looks at convergence every few iterations.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "isdf_internal.h"

#pragma clang fp contract(off)

namespace isdf_hip {

namespace {

constexpr int kSortChunk = 2048;   // points per histogram/scatter chunk
constexpr int kFlagTile = 2048;    // elements per compaction block (256 x 8)
constexpr int kMaxRk = 8192;       // LDS-resident label histograms and sort

__device__ __forceinline__ unsigned long long lanemask_lt() {
  unsigned lane = __lane_id();
  return (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
}

// Reference coordinates, utility.cpp:711-713 (points) and :758-760 (centroids).
__device__ __forceinline__ void grid_coords(int tmp, int n0, int n1, double dx0, double dx1,
                                            double dx2, double* x, double* y, double* z) {
  double gx = (double)(tmp % n0) * dx0;
  double gy = ((double)(tmp % (n0 * n1)) - gx / dx0) / (double)n0 * dx1;
  double gz = (((double)tmp - gx / dx0) - gy / dx1 * (double)n0) / (double)(n0 * n1) * dx2;
  *x = gx;
  *y = gy;
  *z = gz;
}

// Distance of pdist2, utility.cpp:609-612: ((d1*d1 + d2*d2) + d3*d3).
__device__ __forceinline__ double dist2(double px, double py, double pz, double qx, double qy,
                                        double qz) {
  double d1 = px - qx;
  double d2 = py - qy;
  double d3 = pz - qz;
  double s = d1 * d1;
  s = s + d2 * d2;
  s = s + d3 * d3;
  return s;
}

// Block-wide exclusive scan of one int per thread (blockDim.x multiple of 64,
// at most 1024).  Returns the exclusive prefix; *total gets the block sum.
__device__ int block_excl_scan(int v, int* total) {
  __shared__ int wsum[17];
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
  for (int d = 1; d < 64; d <<= 1) {
    int y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  int nw = blockDim.x >> 6;
  if (threadIdx.x == 0) {
    int s = 0;
    for (int i = 0; i < nw; i++) {
      int t = wsum[i];
      wsum[i] = s;
      s += t;
    }
    wsum[16] = s;
  }
  __syncthreads();
  int res = x - v + wsum[wave];
  *total = wsum[16];
  __syncthreads();
  return res;
}

// ---------------------------------------------------------------------------
// Weights, src/cpu/getpoints_Kmeans_MPI.cpp:37-45.
// ---------------------------------------------------------------------------
template <bool BATCH>
__global__ __launch_bounds__(256) void k_weights(const double* __restrict__ psi, long long ld,
                                                 int ng, int lo, int hi, int s1lo, int s1hi,
                                                 int s2lo, int s2hi, double* __restrict__ w,
                                                 const KmJob* jobs) {
  if constexpr (BATCH) {
    const KmJob& J = jobs[blockIdx.y];
    psi = J.psi;
    ld = J.ld;
    ng = J.b.n;
    lo = J.lo;
    hi = J.hi;
    s1lo = J.s1lo;
    s1hi = J.s1hi;
    s2lo = J.s2lo;
    s2hi = J.s2hi;
    w = J.b.w;
  }
  long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ng) return;
  double sI = 0.0, sJ = 0.0;  // sI over S2, sJ over S1 (reference names)
  const double* p = psi + r;
  // 16 states' loads in flight per lane (HBM latency), then the sums in order.
  constexpr int U = 16;
  for (int i0 = lo; i0 < hi; i0 += U) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = (i0 + u < hi) ? p[(long long)(i0 + u) * ld] : 0.0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u;
      if (i < hi) {
        const double q = v[u] * v[u];
        if (i >= s2lo && i < s2hi) sI = sI + q;
        if (i >= s1lo && i < s1hi) sJ = sJ + q;
      }
    }
  }
  double acc = 0.0;
  acc = acc + sI * sJ;
  w[r] = acc;
}

// findMax (numvec_impl.hpp:259-266) as an atomic max over IEEE bit patterns:
// all weights are >= +0, where the bit order is the numeric order.  One
// atomic per workgroup (same-address atomics serialise; the launch keeps to a
// few hundred workgroups).
template <bool BATCH>
__global__ __launch_bounds__(256) void k_maxw(const double* __restrict__ w, int n,
                                              unsigned long long* maxbits, const KmJob* jobs) {
  __shared__ unsigned long long wm[4];
  if constexpr (BATCH) {
    const KmJob& J = jobs[blockIdx.y];
    w = J.b.w;
    n = J.b.n;
    maxbits = (unsigned long long*)(J.b.scal + KS_MAXW);
  }
  unsigned long long m = 0ull;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double v = w[i];
    unsigned long long b = (v > 0.0) ? (unsigned long long)__double_as_longlong(v) : 0ull;
    m = b > m ? b : m;
  }
  for (int d = 32; d > 0; d >>= 1) {
    unsigned long long o = __shfl_xor(m, d);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; k++) m = wm[k] > m ? wm[k] : m;
    atomicMax(maxbits, m);
  }
}

// flags[i] = w[i] > df_tol * maxW (utility.cpp:650-659; every flag set when
// df_tol <= 1e-16), and the number set per tile of kFlagTile into bcount.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_thresh_count(const double* __restrict__ w, int n,
                                                      const unsigned long long* maxbits,
                                                      double df_tol, uint8_t* __restrict__ flags,
                                                      int* __restrict__ bcount,
                                                      const KmJob* jobs) {
  if constexpr (BATCH) {
    const KmJob& J = jobs[blockIdx.y];
    if ((int)blockIdx.x >= J.b.nblk_flag) return;
    w = J.b.w;
    n = J.b.n;
    maxbits = (const unsigned long long*)(J.b.scal + KS_MAXW);
    flags = J.b.flags;
    bcount = J.b.bcount;
  }
  const bool all = !(df_tol > 1e-16);
  const double thr = df_tol * __longlong_as_double((long long)*maxbits);
  const long long i0 = (long long)blockIdx.x * kFlagTile + threadIdx.x * 8;
  int c = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const long long i = i0 + k;
    if (i < n) {
      const bool f = all || w[i] > thr;
      flags[i] = f ? 1 : 0;
      c += f ? 1 : 0;
    }
  }
  int tot;
  block_excl_scan(c, &tot);
  if (threadIdx.x == 0) bcount[blockIdx.x] = tot;
}

// Ordered compaction: count, scan, scatter of indices whose flag == want.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_flag_count(const uint8_t* __restrict__ flags, int n,
                                                    int want, int* __restrict__ bcount,
                                                    const KmJob* jobs) {
  if constexpr (BATCH) {
    const KmJob& J = jobs[blockIdx.y];
    if ((int)blockIdx.x >= J.b.nblk_flag) return;
    flags = J.b.flags;
    n = J.b.n;
    bcount = J.b.bcount;
  }
  long long i0 = (long long)blockIdx.x * kFlagTile + threadIdx.x * 8;
  int c = 0;
  for (int k = 0; k < 8; k++) {
    long long i = i0 + k;
    if (i < n && flags[i] == want) c++;
  }
  int tot;
  block_excl_scan(c, &tot);
  if (threadIdx.x == 0) bcount[blockIdx.x] = tot;
}

// In-place exclusive scan of a[0..m) by one block; a[m] receives the total,
// and so does *total_out when given.
template <bool BATCH>
__global__ __launch_bounds__(1024) void k_scan_single(int* a, int m, int* total_out,
                                                      const KmJob* jobs) {
  __shared__ int carry;
  if constexpr (BATCH) {
    const KmJob& J = jobs[blockIdx.y];
    a = J.b.bcount;
    m = J.b.nblk_flag;
  }
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int off = 0; off < m; off += blockDim.x) {
    int i = off + threadIdx.x;
    int v = (i < m) ? a[i] : 0;
    int tot;
    int ex = block_excl_scan(v, &tot);
    if (i < m) a[i] = ex + carry;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a[m] = carry;
    if (total_out) *total_out = carry;
  }
}

// base_dev, when given, replaces base (an offset still on the device).
template <bool BATCH>
__global__ __launch_bounds__(256) void k_flag_scatter(const uint8_t* __restrict__ flags, int n,
                                                      int want, const int* __restrict__ bexcl,
                                                      int base, int* __restrict__ out,
                                                      const int* base_dev, const KmJob* jobs) {
  if constexpr (BATCH) {  // the tail of piv after the distinct centers
    const KmJob& J = jobs[blockIdx.y];
    if ((int)blockIdx.x >= J.b.nblk_flag) return;
    flags = J.b.flags;
    n = J.b.n;
    bexcl = J.b.bcount;
    out = J.piv;
    base_dev = J.b.scal + KS_RKU;
  }
  if (base_dev) base = *base_dev;
  long long i0 = (long long)blockIdx.x * kFlagTile + threadIdx.x * 8;
  int c = 0;
  for (int k = 0; k < 8; k++) {
    long long i = i0 + k;
    if (i < n && flags[i] == want) c++;
  }
  int tot;
  int pos = base + bexcl[blockIdx.x] + block_excl_scan(c, &tot);
  for (int k = 0; k < 8; k++) {
    long long i = i0 + k;
    if (i < n && flags[i] == want) out[pos++] = (int)i;
  }
}

// Local points, utility.cpp:707-715.  With one rank nptLocal = n; entries of
// `index` past npt read the zero-initialised tail of the reference's vector,
// i.e. grid point 0 (utility.cpp:648, :659, :708).
template <bool BATCH>
__global__ __launch_bounds__(256) void k_setup_points(int n, int npt, const int* __restrict__ index,
                                                      const double* __restrict__ w, int n0, int n1,
                                                      double dx0, double dx1, double dx2,
                                                      int* __restrict__ gidx, double* __restrict__ X,
                                                      double* __restrict__ Y, double* __restrict__ Z,
                                                      double* __restrict__ WL, const KmJob* jobs) {
  if constexpr (BATCH) {
    const KmJob& J = jobs[blockIdx.y];
    n = npt = J.b.n;
    w = J.b.w;
    n0 = J.b.ngrid[0];
    n1 = J.b.ngrid[1];
    dx0 = J.b.dx[0];
    dx1 = J.b.dx[1];
    dx2 = J.b.dx[2];
    gidx = J.b.gidx;
    X = J.b.X;
    Y = J.b.Y;
    Z = J.b.Z;
    WL = J.b.WL;
  }
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int tmp = (npt == n) ? i : ((i < npt) ? index[i] : 0);
  gidx[i] = tmp;
  WL[i] = w[tmp];
  double x, y, z;
  grid_coords(tmp, n0, n1, dx0, dx1, dx2, &x, &y, &z);
  X[i] = x;
  Y[i] = y;
  Z[i] = z;
}

// Initial centroids = the first rk points above the threshold (utility.cpp:750-762).
template <bool BATCH>
__global__ __launch_bounds__(256) void k_init_centroids(int rk, int n, int npt,
                                                        const int* __restrict__ index, int n0,
                                                        int n1, double dx0, double dx1, double dx2,
                                                        double* cx, double* cy, double* cz,
                                                        int* scal, const KmJob* jobs) {
  if constexpr (BATCH) {
    const KmJob& J = jobs[blockIdx.y];
    rk = J.rk;
    n = npt = J.b.n;
    n0 = J.b.ngrid[0];
    n1 = J.b.ngrid[1];
    dx0 = J.b.dx[0];
    dx1 = J.b.dx[1];
    dx2 = J.b.dx[2];
    cx = J.b.cx;
    cy = J.b.cy;
    cz = J.b.cz;
    scal = J.b.scal;
  }
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c == 0) scal[KS_RK] = rk;
  if (c >= rk) return;
  int tmp = (npt == n) ? c : index[c];
  double x, y, z;
  grid_coords(tmp, n0, n1, dx0, dx1, dx2, &x, &y, &z);
  cx[c] = x;
  cy[c] = y;
  cz[c] = z;
}

// Brick order of the grid for the assignment: brick b (8 x 8 x 4 points,
// bricks numbered x fastest), slot s -> grid point or -1 off the grid.  When
// some grid points are below the density threshold the blocks are plain runs
// of 256 local point indices instead.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_build_blocks(int n, int grid_mode, int n0, int n1, int n2,
                                                     int* __restrict__ porder, const KmJob* jobs) {
  if constexpr (BATCH) {
    const KmJob& J = jobs[blockIdx.y];
    if ((int)blockIdx.x >= J.b.nblk_pts) return;
    n = J.b.n;
    grid_mode = 1;
    n0 = J.b.ngrid[0];
    n1 = J.b.ngrid[1];
    n2 = J.b.ngrid[2];
    porder = J.b.porder;
  }
  const int b = blockIdx.x, s = threadIdx.x;
  int v;
  if (grid_mode) {
    const int nbx = (n0 + 7) / 8, nby = (n1 + 7) / 8;
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const int x = bx * 8 + (s & 7), y = by * 8 + ((s >> 3) & 7), z = bz * 4 + (s >> 6);
    v = (x < n0 && y < n1 && z < n2) ? x + n0 * (y + n1 * z) : -1;
  } else {
    const int i = b * 256 + s;
    v = i < n ? i : -1;
  }
  porder[(long long)b * 256 + s] = v;
}

// A value every lane of the wave holds alike, moved to scalar registers.
__device__ __forceinline__ double uniform_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double wave_min(double v) {
  for (int d = 32; d > 0; d >>= 1) v = fmin(v, __shfl_xor(v, d));
  return v;
}

// Minimum of one double per thread over a 256-thread block.
__device__ double block_min256(double v, double* red) {
  v = wave_min(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  const double r = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
  __syncthreads();
  return r;
}

// Candidate-list capacity of the assignment: 256 once the centroids have
// spread; 1024 for the first two assignments, when the initial centroids (the
// first rk grid points, one corner of the grid) leave hundreds of candidates
// per brick.  A brick with more takes its candidates in windows of that size.
constexpr int kCandCap = 256, kCandCapWide = 1024;
// Grid points per thread of a grid-mode assignment brick (8 x 8 x 4 kBrickPZ
// points per workgroup of 256).  Two (one cull for twice the points, a 96^3
// grid in one round of workgroups) measured slower at 64^3 and 96^3: the
// deeper bricks keep more candidates.
constexpr int kBrickPZ = 1;
// Lloyd iterations whose centroid sums use 1024-thread workgroups, 15 producer
// waves (the initial centroids, the first rk grid points, leave clusters of up
// to ~1e5 points for the first few updates); later ones use 320 threads, 4
// producer waves of 2 positions per lane (512-position rounds, 16 KB of LDS:
// more workgroups resident; 64^3 K-means 2.42 -> 2.36 ms).
constexpr int kWideSumsIters = 4;
// Smallest rk for which the bricks cull from super-brick lists: below it a
// brick's own pass over every centroid costs less than the extra launch
// (64^3, rk 1024: 17.5 + 5.8 us per iteration with the lists).
constexpr int kSuperMinRk = 1536;

// Box holding the K-means points of the grid-index ranges [xa, xb] x [ya, yb]
// x [za, zb]: their lattice box widened by 1e-10 of the cell length, which
// covers the few ulps by which the reference's coordinate formula
// (utility.cpp:711-713) departs from the lattice.
__device__ __forceinline__ void brick_box(int xa, int xb, int ya, int yb, int za, int zb, int n0,
                                          int n1, int n2, double dx0, double dx1, double dx2,
                                          double lo[3], double hi[3]) {
  const double m0 = 1e-10 * (dx0 * n0), m1 = 1e-10 * (dx1 * n1), m2 = 1e-10 * (dx2 * n2);
  lo[0] = xa * dx0 - m0;
  hi[0] = xb * dx0 + m0;
  lo[1] = ya * dx1 - m1;
  hi[1] = yb * dx1 + m1;
  lo[2] = za * dx2 - m2;
  hi[2] = zb * dx2 + m2;
}

// Minimum of one double per thread over a block of NT threads.
template <int NT>
__device__ double block_min_nt(double v, double* red) {
  v = wave_min(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int w = 1; w < NT / 64; w++) r = fmin(r, red[w]);
  __syncthreads();
  return r;
}

// Lane-strided culling of the nsrc centroids given by src (list entry idx ->
// centroid index c and its coordinates, in ascending index order) against the
// box [lo, hi]: the kept ones, in the same order, go to out[] (and their
// coordinates to ox / oy / oz when given) in windows of CAP entries, and
// consume(first, count) runs on each window (all threads, after a barrier;
// WINDOWS false: the first window only).  Returns the number kept, which may
// exceed CAP.  A centroid is kept when its lower bound on the distance to
// the box is <= U (1 + 1e-9), U the smallest upper bound over the listed
// centroids.  The coordinates of the first GK chunks of NT entries are read
// once and kept in registers for both passes; later chunks are read again, and
// so are the kept entries for the output (GK is what the caller's occupancy
// affords).  Block-wide (NT
// threads); uses red[NT / 64] and wcnt[32][NT / 64] in LDS (read until the
// last window) and ends with a barrier.
template <int NT, int CAP, int GK, bool WINDOWS, typename SRC, typename CONSUME>
__device__ int cull_box(int nsrc, SRC src, const double lo[3], const double hi[3], int* out,
                        double* ox, double* oy, double* oz, double* red, int (*wcnt)[NT / 64],
                        CONSUME consume) {
  constexpr int NW = NT / 64;
  constexpr int G = 4;
  static_assert(GK >= 1 && GK <= G, "kept chunks");
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int nchunk = (nsrc + NT - 1) / NT;
  auto fetch = [&](int idx, int& c, double& x, double& y, double& z) {
    if (idx < nsrc) {
      src(idx, c, x, y, z);
    } else {
      c = 0;
      x = lo[0];
      y = lo[1];
      z = lo[2];
    }
  };
  auto upper = [&](double x, double y, double z) {
    const double ex = fmax(fabs(x - lo[0]), fabs(x - hi[0]));
    const double ey = fmax(fabs(y - lo[1]), fabs(y - hi[1]));
    const double ez = fmax(fabs(z - lo[2]), fabs(z - hi[2]));
    return ex * ex + ey * ey + ez * ez;
  };
  auto lower = [&](double x, double y, double z) {
    const double ex = fmax(0.0, fmax(lo[0] - x, x - hi[0]));
    const double ey = fmax(0.0, fmax(lo[1] - y, y - hi[1]));
    const double ez = fmax(0.0, fmax(lo[2] - z, z - hi[2]));
    return ex * ex + ey * ey + ez * ez;
  };
  double x0[GK], y0[GK], z0[GK];
  double ub = 1e300;
  {
    double x[G], y[G], z[G];
#pragma unroll
    for (int u = 0; u < G; u++) {
      int c;
      fetch(u * NT + t, c, x[u], y[u], z[u]);
    }
#pragma unroll
    for (int u = 0; u < G; u++)
      if (u * NT + t < nsrc) ub = fmin(ub, upper(x[u], y[u], z[u]));
#pragma unroll
    for (int u = 0; u < GK; u++) {
      x0[u] = x[u];
      y0[u] = y[u];
      z0[u] = z[u];
    }
  }
  for (int j0 = G; j0 < nchunk; j0 += G) {
    int c[G];
    double x[G], y[G], z[G];
#pragma unroll
    for (int u = 0; u < G; u++) fetch((j0 + u) * NT + t, c[u], x[u], y[u], z[u]);
#pragma unroll
    for (int u = 0; u < G; u++)
      if ((j0 + u) * NT + t < nsrc) ub = fmin(ub, upper(x[u], y[u], z[u]));
  }
  const double thr = block_min_nt<NT>(ub, red) * (1.0 + 1e-9);
  unsigned keepbits = 0u;  // bit j: list entry NT j + t is kept (nchunk <= 32)
  auto test = [&](int j, double x, double y, double z) {  // j < nchunk (uniform)
    const bool keep = j * NT + t < nsrc && lower(x, y, z) <= thr;
    const unsigned long long m = __ballot(keep);
    if (keep) keepbits |= 1u << j;
    if (lane == 0) wcnt[j][wave] = __popcll(m);
  };
#pragma unroll
  for (int u = 0; u < GK; u++)
    if (u < nchunk) test(u, x0[u], y0[u], z0[u]);
  for (int j0 = GK; j0 < nchunk; j0 += G) {
    int c[G];
    double x[G], y[G], z[G];
#pragma unroll
    for (int u = 0; u < G; u++) fetch((j0 + u) * NT + t, c[u], x[u], y[u], z[u]);
#pragma unroll
    for (int u = 0; u < G; u++)
      if (j0 + u < nchunk) test(j0 + u, x[u], y[u], z[u]);
  }
  __syncthreads();
  int total = 0;
  for (int j = 0; j < nchunk; j++) {
#pragma unroll
    for (int w2 = 0; w2 < NW; w2++) total += wcnt[j][w2];
  }
  // Kept entries in list order: chunk j before chunk j + 1, inside a chunk by
  // wave, then lane; window [ws, ws + CAP) of them at a time.
  const int nwin = WINDOWS ? (total + CAP - 1) / CAP : 1;
  for (int ws = 0, win = 0; win < nwin; win++, ws += CAP) {
    int nc = 0;
    auto emit = [&](int j) {  // j < nchunk (uniform)
      const bool keep = (keepbits >> j) & 1u;
      const unsigned long long m = __ballot(keep);
      int before = 0, tot = 0;
#pragma unroll
      for (int w2 = 0; w2 < NW; w2++) {
        const int v = wcnt[j][w2];
        before += (w2 < wave) ? v : 0;
        tot += v;
      }
      if (keep) {
        const int pos = nc + before + __popcll(m & lanemask_lt()) - ws;
        if (pos >= 0 && pos < CAP) {
          int c;
          double x, y, z;
          src(j * NT + t, c, x, y, z);
          out[pos] = c;
          if (ox) {
            ox[pos] = x;
            oy[pos] = y;
            oz[pos] = z;
          }
        }
      }
      nc += tot;
    };
    for (int j = 0; j < nchunk; j++) emit(j);
    __syncthreads();
    consume(ws, min(CAP, total - ws));
    if (win + 1 < nwin) __syncthreads();  // the window is read before the next is written
  }
  return total;
}

// Candidate lists of the super-bricks (2 x 2 x 2 bricks, 16 x 16 x 8 grid
// points) against every centroid, for k_assign_culled to cull its brick from:
// the centroid indices (slist) and, in the same order, their coordinates
// (slxyz: x, y, z planes of nsb * kSuperCap each).
// A centroid dropped here is, for every point of the super-brick, farther
// than the nearest centroid by a relative 1e-9, so it is none of its bricks'
// candidates either.  Grid mode (every grid point a K-means point) only.
__global__ __launch_bounds__(1024) void k_cull_super(const double* __restrict__ cx,
                                                     const double* __restrict__ cy,
                                                     const double* __restrict__ cz, int n0,
                                                     int n1, int n2, double dx0, double dx1,
                                                     double dx2, int* __restrict__ slist,
                                                     double* __restrict__ slxyz,
                                                     int* __restrict__ scount, const int* scal) {
  __shared__ double red[16];
  __shared__ int wcnt[32][16];
  if (scal[KS_ACTIVE] == 0) return;
  const int rk = scal[KS_RK];
  const int t = threadIdx.x;
  const int nsbx = (n0 + 15) / 16, nsby = (n1 + 15) / 16;
  const int sb = blockIdx.x;
  const int sx0 = (sb % nsbx) * 16, sy0 = ((sb / nsbx) % nsby) * 16, sz0 = (sb / (nsbx * nsby)) * 8;
  double lo[3], hi[3];
  brick_box(sx0, min(sx0 + 15, n0 - 1), sy0, min(sy0 + 15, n1 - 1), sz0, min(sz0 + 7, n2 - 1), n0,
            n1, n2, dx0, dx1, dx2, lo, hi);
  const long long o = (long long)sb * kSuperCap, plane = (long long)gridDim.x * kSuperCap;
  auto all = [=](int idx, int& c, double& x, double& y, double& z) {
    c = idx;
    x = cx[idx];
    y = cy[idx];
    z = cz[idx];
  };
  // One window: a list that overflows is not used (its bricks cull from all).
  const int nc = cull_box<1024, kSuperCap, 2, false>(rk, all, lo, hi, slist + o, slxyz + o,
                                                  slxyz + plane + o, slxyz + 2 * plane + o, red,
                                                  wcnt, [](int, int) {});
  if (t == 0) scount[sb] = nc;
}

// Fused pdist2 + findMin(Dim=1) (utility.cpp:598-615, :535-562) with
// conservative culling.  A workgroup takes one block of points, bounds their
// coordinates by a box, and keeps only the centroids whose lower bound LB on
// the distance to the box is <= U (1 + 1e-9), U the smallest upper bound over
// all centroids.  A centroid left out is, for every point of the block,
// farther than the nearest centroid by a relative 1e-9, far beyond the few
// ulps of rounding in the distance, so it can neither win nor tie; the first
// minimum over the candidates, kept in ascending order, is therefore the
// reference's first minimum over all centroids.
// Blocks: in grid mode a brick of 8 x 8 x 4 PZ grid points, thread t holding
// points (t & 7, (t >> 3) & 7, (t >> 6) + 4 p) of it for p < PZ (one cull
// serves PZ x 256 points); otherwise (PZ = 1) a run of 256 entries of the
// point order.  In grid mode the workgroup also folds its points into the
// bounding boxes (grid indices) of their clusters, which the centroid sums
// walk.
template <bool BATCH, int CAP, int PZ>
__global__ __launch_bounds__(256) void k_assign_culled(
    const int* __restrict__ porder, const double* __restrict__ X, const double* __restrict__ Y,
    const double* __restrict__ Z, const double* __restrict__ cx, const double* __restrict__ cy,
    const double* __restrict__ cz, int* __restrict__ label, int* scal, int count_changes,
    int* __restrict__ bbox, int rkmax, int grid_mode, int n0, int n1, int n2, double dx0,
    double dx1, double dx2, int* __restrict__ changed_blk, const int* __restrict__ slist,
    const double* __restrict__ slxyz, const int* __restrict__ scount, int nsb,
    const KmJob* jobs) {
  __shared__ int cand[CAP];
  int b = blockIdx.x;  // the block (brick) of the instance
  if constexpr (BATCH) {
    // porder is the batch's brick list: (job, brick) of block x at
    // porder[2x], porder[2x + 1]; only bricks of the instances are launched.
    const int2 jb = *reinterpret_cast<const int2*>(porder + 2 * (long long)blockIdx.x);
    const KmJob& J = jobs[jb.x];
    b = jb.y;
    porder = J.b.porder;
    X = J.b.X;
    Y = J.b.Y;
    Z = J.b.Z;
    cx = J.b.cx;
    cy = J.b.cy;
    cz = J.b.cz;
    label = J.b.label;
    scal = J.b.scal;
    bbox = J.b.bbox;
    rkmax = J.b.rkmax;
    grid_mode = 1;
    n0 = J.b.ngrid[0];
    n1 = J.b.ngrid[1];
    n2 = J.b.ngrid[2];
    dx0 = J.b.dx[0];
    dx1 = J.b.dx[1];
    dx2 = J.b.dx[2];
    changed_blk = J.b.changed_blk;
    slist = nullptr;
  }
  // Candidate coordinates; once the distances are done the same LDS holds the
  // block's box of each candidate's points (6 ints per candidate, cbox).
  __shared__ __align__(16) double ccs[3 * CAP];
  double* const ccx = ccs;
  double* const ccy = ccs + CAP;
  double* const ccz = ccs + 2 * CAP;
  int(*const cbox)[CAP] = reinterpret_cast<int(*)[CAP]>(ccs);
  __shared__ double red[8];
  __shared__ double rbox[6][4];
  if (count_changes && scal[KS_ACTIVE] == 0) return;
  const int rk = scal[KS_RK];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int nbx = (n0 + 7) / 8, nby = (n1 + 7) / 8;
  const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);  // grid mode
  int i[PZ];
  if (grid_mode) {
    const int x = bx * 8 + (t & 7), y = by * 8 + ((t >> 3) & 7);
#pragma unroll
    for (int p = 0; p < PZ; p++) {
      const int z = bz * 4 * PZ + (t >> 6) + 4 * p;
      i[p] = (x < n0 && y < n1 && z < n2) ? x + n0 * (y + n1 * z) : -1;
    }
  } else {
    i[0] = porder[(long long)b * 256 + t];  // PZ == 1 here
  }
  bool valid[PZ];
  double px[PZ], py[PZ], pz[PZ];
  int old[PZ];
#pragma unroll
  for (int p = 0; p < PZ; p++) {
    valid[p] = i[p] >= 0;
    px[p] = py[p] = pz[p] = 0.0;
    if (valid[p]) {
      if (grid_mode) {
        // Point i is grid point i: its coordinates from the same formula and
        // operations as k_setup_points (bit for bit what X, Y, Z hold), so no
        // load waits here; the old label is needed only at the end.
        grid_coords(i[p], n0, n1, dx0, dx1, dx2, &px[p], &py[p], &pz[p]);
      } else {
        px[p] = X[i[p]];
        py[p] = Y[i[p]];
        pz[p] = Z[i[p]];
      }
    }
    old[p] = (count_changes && valid[p]) ? label[i[p]] : -1;
  }

  double lo[3], hi[3];
  if (grid_mode) {
    // The box of the brick from its grid-index ranges: the coordinate formula
    // (utility.cpp:711-713) puts every point within a few ulps of its lattice
    // position, so the lattice box widened by 1e-10 L holds every point (a
    // wider box only keeps more candidates).  A brick always holds a point.
    brick_box(bx * 8, min(bx * 8 + 7, n0 - 1), by * 8, min(by * 8 + 7, n1 - 1), bz * 4 * PZ,
              min(bz * 4 * PZ + 4 * PZ - 1, n2 - 1), n0, n1, n2, dx0, dx1, dx2, lo, hi);
  } else {
    // Bounding box of the block's points.
    double v[6] = {valid[0] ? px[0] : 1e300,  valid[0] ? py[0] : 1e300,
                   valid[0] ? pz[0] : 1e300,  valid[0] ? -px[0] : 1e300,
                   valid[0] ? -py[0] : 1e300, valid[0] ? -pz[0] : 1e300};
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const double m = wave_min(v[k]);
      if (lane == 0) rbox[k][wave] = m;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = fmin(fmin(rbox[k][0], rbox[k][1]), fmin(rbox[k][2], rbox[k][3]));
      hi[k] = -fmin(fmin(rbox[k + 3][0], rbox[k + 3][1]), fmin(rbox[k + 3][2], rbox[k + 3][3]));
    }
    if (lo[0] > 1e299) {  // no point in this block (uniform)
      if (count_changes && t == 0) changed_blk[b] = 0;
      return;
    }
  }

  // The box is uniform over the workgroup: scalar registers.
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = uniform_f64(lo[k]);
    hi[k] = uniform_f64(hi[k]);
  }

  // The centroids to cull: all of them, or (one instance, grid mode) the list
  // of the brick's super-brick (16 x 16 x 8 grid points), when it did not
  // overflow.
  const int* sl = nullptr;
  const double *slx = nullptr, *sly = nullptr, *slz = nullptr;
  int nsrc = rk;
  if (slist) {
    const int nsbx = (n0 + 15) / 16, nsby = (n1 + 15) / 16;
    const int sb = (bx >> 1) + nsbx * ((by >> 1) + nsby * ((bz * PZ) >> 1));
    const int ns = scount[sb];
    if (ns <= kSuperCap) {
      const long long o = (long long)sb * kSuperCap, plane = (long long)nsb * kSuperCap;
      sl = slist + o;
      slx = slxyz + o;
      sly = slxyz + plane + o;
      slz = slxyz + 2 * plane + o;
      nsrc = ns;
    }
  }
  __shared__ int wcnt[32][4];
  auto src = [=](int idx, int& c, double& x, double& y, double& z) {
    if (sl) {
      c = sl[idx];
      x = slx[idx];
      y = sly[idx];
      z = slz[idx];
    } else {
      c = idx;
      x = cx[idx];
      y = cy[idx];
      z = cz[idx];
    }
  };
  // First minimum over the candidates (ascending), window by window, for the
  // PZ points at once: bi the centroid, bk its place in the candidate list.
  double best[PZ];
  int bk[PZ], bi[PZ];
#pragma unroll
  for (int p = 0; p < PZ; p++) {
    best[p] = __builtin_inf();
    bk[p] = 0;
    bi[p] = 0;
  }
  auto consume = [&](int ws, int n) {
    int kw[PZ];
#pragma unroll
    for (int p = 0; p < PZ; p++) kw[p] = -1;
    int k = 0;
    for (; k + 1 < n; k += 2) {
      const double qx0 = ccx[k], qy0 = ccy[k], qz0 = ccz[k];
      const double qx1 = ccx[k + 1], qy1 = ccy[k + 1], qz1 = ccz[k + 1];
#pragma unroll
      for (int p = 0; p < PZ; p++) {
        const double d0 = dist2(px[p], py[p], pz[p], qx0, qy0, qz0);
        const double d1 = dist2(px[p], py[p], pz[p], qx1, qy1, qz1);
        if (d0 < best[p]) { best[p] = d0; kw[p] = k; }
        if (d1 < best[p]) { best[p] = d1; kw[p] = k + 1; }
      }
    }
    if (k < n) {
      const double qx = ccx[k], qy = ccy[k], qz = ccz[k];
#pragma unroll
      for (int p = 0; p < PZ; p++) {
        const double d = dist2(px[p], py[p], pz[p], qx, qy, qz);
        if (d < best[p]) { best[p] = d; kw[p] = k; }
      }
    }
#pragma unroll
    for (int p = 0; p < PZ; p++) {
      if (kw[p] >= 0) {
        bk[p] = ws + kw[p];
        bi[p] = cand[kw[p]];
      }
    }
  };
  const int nc = cull_box<256, CAP, (PZ == 1 ? 2 : 1), true>(nsrc, src, lo, hi, cand, ccx, ccy,
                                                             ccz, red, wcnt, consume);
  // One window: the block's cluster boxes fold through LDS (cbox, by place in
  // the candidate list); several: straight into the global boxes.
  const bool culled = nc <= CAP;
  __syncthreads();  // every thread is done with the candidates' coordinates
  if (culled && grid_mode) {
    for (int k = t; k < nc; k += 256) {
      cbox[0][k] = 0x7fffffff;
      cbox[1][k] = 0x7fffffff;
      cbox[2][k] = 0x7fffffff;
      cbox[3][k] = -1;
      cbox[4][k] = -1;
      cbox[5][k] = -1;
    }
  }
  __syncthreads();
  int nchanged = 0;
#pragma unroll
  for (int p = 0; p < PZ; p++) {
    if (valid[p]) {
      if (count_changes) nchanged += (old[p] != bi[p]) ? 1 : 0;
      label[i[p]] = bi[p];
    }
  }
  if (count_changes) {
    // Per-block count; the loop control sums the blocks (no contended atomics).
    int v = nchanged;
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
    if (lane == 0) wcnt[0][wave] = v;
    __syncthreads();
    if (t == 0) changed_blk[b] = wcnt[0][0] + wcnt[0][1] + wcnt[0][2] + wcnt[0][3];
  }
  if (!grid_mode) return;
  if (culled) {
    // Boxes of the clusters, per wave and z plane: lane l holds brick point
    // (x, y) = (l & 7, l >> 3) of the plane, so the ballot of one label's lanes
    // gives its x and y extent by bit folding (uniform, scalar work) and one
    // lane folds it into the block's box of that candidate: no contended LDS
    // atomics.
    const int bx0 = bx * 8, by0 = by * 8;
#pragma unroll
    for (int p = 0; p < PZ; p++) {
      const int zw = bz * 4 * PZ + wave + 4 * p;
      unsigned long long todo = __ballot(valid[p]);
      while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int kk = __shfl(bk[p], leader);
        const unsigned long long m = __ballot(valid[p] && bk[p] == kk);
        todo &= ~m;
        unsigned cols = 0u, rows = 0u;
#pragma unroll
        for (int yy = 0; yy < 8; yy++) {
          const unsigned r8 = (unsigned)(m >> (8 * yy)) & 0xffu;
          cols |= r8;
          rows |= (r8 ? 1u : 0u) << yy;
        }
        if (lane == leader) {
          atomicMin(&cbox[0][kk], bx0 + __ffs(cols) - 1);
          atomicMin(&cbox[1][kk], by0 + __ffs(rows) - 1);
          atomicMin(&cbox[2][kk], zw);
          atomicMax(&cbox[3][kk], bx0 + 31 - __clz(cols));
          atomicMax(&cbox[4][kk], by0 + 31 - __clz(rows));
          atomicMax(&cbox[5][kk], zw);
        }
      }
    }
    __syncthreads();
    for (int k = t; k < nc; k += 256) {
      if (cbox[3][k] < 0) continue;
      const int c = cand[k];
      atomicMin(&bbox[c], cbox[0][k]);
      atomicMin(&bbox[rkmax + c], cbox[1][k]);
      atomicMin(&bbox[2 * rkmax + c], cbox[2][k]);
      atomicMax(&bbox[3 * rkmax + c], cbox[3][k]);
      atomicMax(&bbox[4 * rkmax + c], cbox[4][k]);
      atomicMax(&bbox[5 * rkmax + c], cbox[5][k]);
    }
  } else {
#pragma unroll
    for (int p = 0; p < PZ; p++) {
      if (!valid[p]) continue;
      const int q = bi[p];
      const int gx = i[p] % n0, gy = (i[p] / n0) % n1, gz = i[p] / (n0 * n1);
      atomicMin(&bbox[q], gx);
      atomicMin(&bbox[rkmax + q], gy);
      atomicMin(&bbox[2 * rkmax + q], gz);
      atomicMax(&bbox[3 * rkmax + q], gx);
      atomicMax(&bbox[4 * rkmax + q], gy);
      atomicMax(&bbox[5 * rkmax + q], gz);
    }
  }
}

template <bool BATCH>
__global__ void k_reset_boxes(int* bbox, int rkmax, const KmJob* jobs) {
  if constexpr (BATCH) {
    bbox = jobs[blockIdx.y].b.bbox;
    rkmax = jobs[blockIdx.y].b.rkmax;
  }
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < 6 * rkmax; e += gridDim.x * blockDim.x)
    bbox[e] = (e < 3 * rkmax) ? 0x7fffffff : -1;
}

// Position of this thread's first member among a round's members, taken in
// position order (thread-major, slots q = 0..PT-1 of a thread in order), and
// the round's total, with one barrier: ballots rank the members inside a wave,
// LDS (wsum[NT / 64]) carries the wave totals.  Needs a barrier between the
// reads of wsum here and the next call's writes.
template <int NT, int PT>
__device__ int rank_members(const bool* mem, int* total, int* wsum) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int below = 0, wtot = 0;
#pragma unroll
  for (int q = 0; q < PT; q++) {
    const unsigned long long m = __ballot(mem[q]);
    below += __popcll(m & lanemask_lt());
    wtot += __popcll(m);
  }
  if (lane == 0) wsum[wave] = wtot;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NW; w++) {
    const int v = wsum[w];
    off += (w < wave) ? v : 0;
    tot += v;
  }
  *total = tot;
  return off + below;
}

// 32 steps of four sequential FP64 sums in one wave: row q (lanes 16q..16q+15)
// carries sum q, and lane 16q + k holds its terms 2k (lo) and 2k+1 (hi) of
// this batch.  Step m broadcasts lane m/2 of every row (DPP row_newbcast) into
// acc = fma(term, 1.0, acc), which rounds exactly as acc + term: every lane of
// a row ends with its sum, one instruction per step for all four sums.  The
// leading s_nop covers the DPP hazards (a VALU write of a term or of EXEC just
// before the block).
__device__ __forceinline__ void add32_rows(double& acc, double lo, double hi, double one) {
#define ISDF_BCAST2(K)                                                      \
  "v_fmac_f64_dpp %0, %1, %3 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n" \
  "v_fmac_f64_dpp %0, %2, %3 row_newbcast:" #K " row_mask:0xf bank_mask:0xf\n"
  __asm__ volatile("s_nop 4\n" ISDF_BCAST2(0) ISDF_BCAST2(1) ISDF_BCAST2(2) ISDF_BCAST2(3)
                   ISDF_BCAST2(4) ISDF_BCAST2(5) ISDF_BCAST2(6) ISDF_BCAST2(7) ISDF_BCAST2(8)
                   ISDF_BCAST2(9) ISDF_BCAST2(10) ISDF_BCAST2(11) ISDF_BCAST2(12)
                   ISDF_BCAST2(13) ISDF_BCAST2(14) ISDF_BCAST2(15)
                   : "+v"(acc)
                   : "v"(lo), "v"(hi), "v"(one));
#undef ISDF_BCAST2
}

// Centroid sums, utility.cpp:816-822, when every grid point is a K-means point
// (local index == grid index): one workgroup per cluster walks the cluster's
// bounding box in ascending grid order (z, y, x) and adds the terms w, w x,
// w y, w z of the points carrying its label in that order: the reference's
// sequence of roundings (it adds in point order) without sorting.
// Wave 0 is the adder (add32_rows: the four sums in its four rows); waves
// 1..NP produce: a round of CH = NP 64 PT box positions, producer wave pw
// taking positions pw 64 PT .. of it, PT consecutive ones per lane.  While the
// adder adds round r from LDS, the producers load and rank round r + 1 (labels
// two rounds ahead, member terms one round ahead); two barriers per round
// hand the buffer over.
template <bool BATCH, int NP, int PT>
__global__ __launch_bounds__(64 * (NP + 1)) void k_centroid_sums_box(
    int n0, int n1, const int* __restrict__ label, const double* __restrict__ WL,
    const double* __restrict__ X, const double* __restrict__ Y, const double* __restrict__ Z,
    const int* __restrict__ bbox, int rkmax, double* __restrict__ sw, double* __restrict__ sx,
    double* __restrict__ sy, double* __restrict__ sz, const int* scal, int* __restrict__ cost,
    const int* __restrict__ order, const KmJob* jobs) {
  constexpr int CH = NP * 64 * PT;
  static_assert(CH % 32 == 0, "rounds are whole batches of 32 terms");
  // Row q: term q (w, w x, w y, w z) of the round's members in grid order, then
  // zeros up to the next multiple of 32 (the adder takes whole batches).
  __shared__ __align__(16) double buf[4][CH];
  __shared__ __align__(16) int wsum[NP];
  int cl = blockIdx.x;  // the workgroup's place in the cluster order
  if constexpr (BATCH) {
    // order is the batch's cluster list: (job, cluster) of block x at
    // order[2x], order[2x + 1] (every cluster of every instance, in index order).
    const int2 jc = *reinterpret_cast<const int2*>(order + 2 * (long long)blockIdx.x);
    const KmJob& J = jobs[jc.x];
    cl = jc.y;
    n0 = J.b.ngrid[0];
    n1 = J.b.ngrid[1];
    label = J.b.label;
    WL = J.b.WL;
    X = J.b.X;
    Y = J.b.Y;
    Z = J.b.Z;
    bbox = J.b.bbox;
    rkmax = J.b.rkmax;
    sw = J.b.sw;
    sx = J.b.sx;
    sy = J.b.sy;
    sz = J.b.sz;
    scal = J.b.scal;
    cost = nullptr;
    order = nullptr;
  }
  if (scal[KS_ACTIVE] == 0) return;
  const int rk = scal[KS_RK];
  if (cl >= rk) return;
  // Workgroup -> cluster: costliest first when an order is given.
  const int c = order ? order[cl] : cl;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const bool adder = wave == 0;
  const int pw = wave - 1;              // producer wave (adder: -1)
  const int pt = t - 64;                // producer thread (adder: < 0)
  const int x0 = bbox[c], y0 = bbox[rkmax + c], z0 = bbox[2 * rkmax + c];
  const int x1 = bbox[3 * rkmax + c], y1 = bbox[4 * rkmax + c], z1 = bbox[5 * rkmax + c];
  double acc = 0.0;  // the adder's sums, one per row
  // Cost of this cluster for the dispatch order of the next sums: box
  // positions plus twice the members (a member costs about two positions).
  int box_cost = 0;
  if (x1 >= x0) {
    const int bw = x1 - x0 + 1, bh = y1 - y0 + 1;
    const int plane = bw * bh;  // box positions < Ng < 2^31
    const int total = plane * (z1 - z0 + 1);
    const int nround = (total + CH - 1) / CH;
    box_cost = total;
    // Labels of round r (producers): positions r CH + PT pt .. + PT - 1 in
    // grid order ((x, y, z) of the first from one division pair, then
    // stepping); g = -1 past the box.  The loads stay in flight until used.
    auto load_labels = [&](int r, int* g, int* l) {
      const int j = r * CH + PT * pt;
      int z = j / plane, rr = j - z * plane;
      int y = rr / bw, x = rr - y * bw;
#pragma unroll
      for (int q = 0; q < PT; q++) {
        g[q] = -1;
        l[q] = -1;
        if (j + q < total) {
          g[q] = (x0 + x) + n0 * ((y0 + y) + n1 * (z0 + z));
          l[q] = label[g[q]];
        }
        if (++x == bw) {
          x = 0;
          if (++y == bh) {
            y = 0;
            ++z;
          }
        }
      }
    };
    // Members of a round and the loads of their terms (w, x, y, z).
    double tw[PT], tx[PT], ty[PT], tz[PT];
    bool mem[PT];
    auto gather = [&](const int* g, const int* l) {
#pragma unroll
      for (int q = 0; q < PT; q++) {
        mem[q] = l[q] == c;
        if (mem[q]) {
          tw[q] = WL[g[q]];
          tx[q] = X[g[q]];
          ty[q] = Y[g[q]];
          tz[q] = Z[g[q]];
        }
      }
    };
    // Rank inside the producer wave (ballots; slots in position order) and
    // the wave's count to LDS; the offset over waves is read after a barrier.
    int below = 0;
    auto rank = [&]() {
      int wtot = 0;
      below = 0;
#pragma unroll
      for (int q = 0; q < PT; q++) {
        const unsigned long long m = __ballot(mem[q]);
        below += __popcll(m & lanemask_lt());
        wtot += __popcll(m);
      }
      if (lane == 0) wsum[pw] = wtot;
    };
    auto round_total = [&](int* off) {  // after the barrier that follows rank()
      int o = 0, tot = 0;
#pragma unroll
      for (int w = 0; w < NP; w++) {
        const int v = wsum[w];
        o += (w < pw) ? v : 0;
        tot += v;
      }
      *off = o;
      return tot;
    };
    // Members to LDS in position order; products w*x (one rounding each, as
    // the reference forms them).  Then +0.0 up to the next multiple of 32 (at
    // most CH): a sum that starts at +0.0 never becomes -0.0, so adding +0.0
    // leaves it exactly as it is.
    auto store = [&](int pos, int tot_round) {
#pragma unroll
      for (int q = 0; q < PT; q++) {
        if (mem[q]) {
          buf[0][pos] = tw[q];
          buf[1][pos] = tw[q] * tx[q];
          buf[2][pos] = tw[q] * ty[q];
          buf[3][pos] = tw[q] * tz[q];
          pos++;
        }
      }
      if (pt < 128) {
        const int i = tot_round + (pt & 31);
        if (i < ((tot_round + 31) & ~31)) buf[pt >> 5][i] = 0.0;
      }
    };
    // The adder: one round's members in order, batches of 32 terms, each one
    // LDS read per lane (two terms of its row) and 32 DPP steps; the reads run
    // two batches ahead.
    const double one = 1.0;
    auto add_round = [&](int tot_round) {
      const int nb = (tot_round + 31) >> 5;
      if (nb == 0) return;
      const double* row = &buf[lane >> 4][2 * (lane & 15)];
      // Batch j (clamped: past the last batch it is read again, not added).
      auto rd = [&](int j) { return *(const double2*)(row + 32 * min(j, nb - 1)); };
      // Three register sets in rotation (unrolled, so no copies between them).
      double2 a = rd(0), b = rd(1), cc;
      for (int j = 0;; j += 3) {
        cc = rd(j + 2);
        add32_rows(acc, a.x, a.y, one);
        if (j + 1 >= nb) break;
        a = rd(j + 3);
        add32_rows(acc, b.x, b.y, one);
        if (j + 2 >= nb) break;
        b = rd(j + 4);
        add32_rows(acc, cc.x, cc.y, one);
        if (j + 3 >= nb) break;
      }
    };
    int gn[PT], ln[PT], gf[PT], lf[PT];  // producers: round r+1 (next), round r+2 (far)
    int tot = 0;
    if (!adder) {
      load_labels(0, gn, ln);
      if (nround > 1) load_labels(1, gf, lf);
      gather(gn, ln);
      rank();
    }
    __syncthreads();
    {
      int off;
      tot = round_total(&off);
      if (!adder) store(off + below, tot);
    }
    __syncthreads();
    for (int r = 0; r < nround; r++) {
      const int tot_r = tot;
      box_cost += 2 * tot_r;
      if (adder) {
        add_round(tot_r);
      } else if (r + 1 < nround) {
#pragma unroll
        for (int q = 0; q < PT; q++) {
          gn[q] = gf[q];
          ln[q] = lf[q];
        }
        if (r + 2 < nround) load_labels(r + 2, gf, lf);
        gather(gn, ln);
        rank();
      }
      __syncthreads();  // the adder is done with round r; round r+1 is ranked
      if (r + 1 < nround) {
        int off;
        tot = round_total(&off);
        if (!adder) store(off + below, tot);
        __syncthreads();  // round r+1 is in the buffer
      }
    }
  }
  if (t == 0 && cost) cost[c] = box_cost;
  if (t == 0) sw[c] = acc;
  if (t == 16) sx[c] = acc;
  if (t == 32) sy[c] = acc;
  if (t == 48) sz[c] = acc;
}

// findMin(Dim=0) (utility.cpp:926-945) when every grid point is a K-means
// point.  The grid point nearest in each coordinate is within dx/2, so the
// nearest point of centroid c lies within sqrt(3)/2 max(dx) of it in every
// coordinate, and so does any point at the same distance: one wave scans that
// window (plus one cell for rounding) and keeps the lexicographic minimum of
// (distance, grid index), i.e. the first minimum in grid order.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_final_window(int n0, int n1, int n2, double dx0,
                                                      double dx1, double dx2,
                                                      const double* __restrict__ X,
                                                      const double* __restrict__ Y,
                                                      const double* __restrict__ Z,
                                                      const double* __restrict__ cx,
                                                      const double* __restrict__ cy,
                                                      const double* __restrict__ cz,
                                                      int* __restrict__ imin, const int* scal,
                                                      const KmJob* jobs) {
  if constexpr (BATCH) {
    const KmJob& J = jobs[blockIdx.y];
    n0 = J.b.ngrid[0];
    n1 = J.b.ngrid[1];
    n2 = J.b.ngrid[2];
    dx0 = J.b.dx[0];
    dx1 = J.b.dx[1];
    dx2 = J.b.dx[2];
    X = J.b.X;
    Y = J.b.Y;
    Z = J.b.Z;
    cx = J.b.cx;
    cy = J.b.cy;
    cz = J.b.cz;
    imin = J.b.imin;
    scal = J.b.scal;
  }
  const int rk = scal[KS_RK];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + wave;
  if (c >= rk) return;
  const double qx = cx[c], qy = cy[c], qz = cz[c];
  const double rho = 0.8660254037844387 * fmax(dx0, fmax(dx1, dx2)) * (1.0 + 1e-9);
  const int x0 = max(0, (int)floor((qx - rho) / dx0) - 1), x1 = min(n0 - 1, (int)ceil((qx + rho) / dx0) + 1);
  const int y0 = max(0, (int)floor((qy - rho) / dx1) - 1), y1 = min(n1 - 1, (int)ceil((qy + rho) / dx1) + 1);
  const int z0 = max(0, (int)floor((qz - rho) / dx2) - 1), z1 = min(n2 - 1, (int)ceil((qz + rho) / dx2) + 1);
  const int bw = max(0, x1 - x0 + 1), bh = max(0, y1 - y0 + 1), bd = max(0, z1 - z0 + 1);
  const int total = bw * bh * bd;
  double best = 1e300;
  int bg = 0x7fffffff;
  for (int j = lane; j < total; j += 64) {
    const int x = x0 + j % bw, y = y0 + (j / bw) % bh, z = z0 + j / (bw * bh);
    const int g = x + n0 * (y + n1 * z);
    const double d = dist2(X[g], Y[g], Z[g], qx, qy, qz);
    if (d < best || (d == best && g < bg)) {
      best = d;
      bg = g;
    }
  }
  for (int dd = 32; dd > 0; dd >>= 1) {
    const double ob = __shfl_xor(best, dd);
    const int og = __shfl_xor(bg, dd);
    if (ob < best || (ob == best && og < bg)) {
      best = ob;
      bg = og;
    }
  }
  if (lane == 0) imin[c] = bg;
}

// Stable counting sort of point indices by label: chunk histograms.
__global__ __launch_bounds__(256) void k_sort_hist(int n, const int* __restrict__ label,
                                                   int* __restrict__ hist, const int* scal) {
  extern __shared__ int h[];
  if (scal[KS_ACTIVE] == 0) return;
  const int rk = scal[KS_RK];
  for (int c = threadIdx.x; c < rk; c += blockDim.x) h[c] = 0;
  __syncthreads();
  int i0 = blockIdx.x * kSortChunk, i1 = min(n, i0 + kSortChunk);
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) atomicAdd(&h[label[i]], 1);
  __syncthreads();
  int* out = hist + (long long)blockIdx.x * rk;
  for (int c = threadIdx.x; c < rk; c += blockDim.x) out[c] = h[c];
}

// Per-cluster exclusive scan over chunks (chunk-major layout); totals -> cnt_pts.
__global__ __launch_bounds__(256) void k_sort_scan_chunks(int nchunk, int* __restrict__ hist,
                                                          int* __restrict__ cnt_pts,
                                                          const int* scal) {
  if (scal[KS_ACTIVE] == 0) return;
  const int rk = scal[KS_RK];
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= rk) return;
  int s = 0;
  for (int b = 0; b < nchunk; b++) {
    int* p = hist + (long long)b * rk + c;
    int t = *p;
    *p = s;
    s += t;
  }
  cnt_pts[c] = s;
}

// Exclusive scan of the cluster sizes -> cluster start offsets (one block).
__global__ __launch_bounds__(1024) void k_sort_scan_clusters(const int* __restrict__ cnt_pts,
                                                             int* __restrict__ base,
                                                             const int* scal) {
  __shared__ int carry;
  if (scal[KS_ACTIVE] == 0) return;
  const int rk = scal[KS_RK];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int off = 0; off < rk; off += blockDim.x) {
    int c = off + threadIdx.x;
    int v = (c < rk) ? cnt_pts[c] : 0;
    int tot;
    int ex = block_excl_scan(v, &tot);
    if (c < rk) base[c] = ex + carry;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
}

// Stable scatter: one wave per chunk walks its points in order; lanes with the
// same label are ranked with ballots, so equal labels keep ascending order.
__global__ __launch_bounds__(64) void k_sort_scatter(int n, const int* __restrict__ label,
                                                     const int* __restrict__ hist,
                                                     const int* __restrict__ base,
                                                     int* __restrict__ sorted, const int* scal) {
  extern __shared__ int run[];
  if (scal[KS_ACTIVE] == 0) return;
  const int rk = scal[KS_RK];
  const int b = blockIdx.x;
  const int* hx = hist + (long long)b * rk;
  for (int c = threadIdx.x; c < rk; c += 64) run[c] = base[c] + hx[c];
  __syncthreads();
  const unsigned long long below = lanemask_lt();
  int i0 = b * kSortChunk, i1 = min(n, i0 + kSortChunk);
  for (int s = i0; s < i1; s += 64) {
    int i = s + (int)threadIdx.x;
    bool valid = i < i1;
    int L = valid ? label[i] : -1;
    unsigned long long todo = __ballot(valid);
    int pos = 0;
    while (todo) {
      int leader = __ffsll((long long)todo) - 1;
      int Lr = __shfl(L, leader);
      bool mine = valid && (L == Lr);
      unsigned long long m = __ballot(mine);
      int r0 = run[Lr];
      if (mine) pos = r0 + __popcll(m & below);
      __builtin_amdgcn_wave_barrier();
      if ((int)threadIdx.x == leader) run[Lr] = r0 + __popcll(m);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      todo &= ~m;
    }
    if (valid) sorted[pos] = i;
  }
}

// Centroid sums, utility.cpp:816-822: one wave per cluster; members in
// ascending point order; lanes 0..3 carry the four sequential sums.
__global__ __launch_bounds__(256) void k_centroid_sums(const int* __restrict__ sorted,
                                                       const int* __restrict__ base,
                                                       const int* __restrict__ cnt_pts,
                                                       const double* __restrict__ WL,
                                                       const double* __restrict__ X,
                                                       const double* __restrict__ Y,
                                                       const double* __restrict__ Z,
                                                       double* __restrict__ sw,
                                                       double* __restrict__ sx,
                                                       double* __restrict__ sy,
                                                       double* __restrict__ sz, const int* scal) {
  __shared__ double buf[4][4][64];
  if (scal[KS_ACTIVE] == 0) return;
  const int rk = scal[KS_RK];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + wave;
  if (c >= rk) return;
  const int s0 = base[c], m = cnt_pts[c];
  double acc = 0.0;
  for (int off = 0; off < m; off += 64) {
    int k = off + lane;
    if (k < m) {
      int i = sorted[s0 + k];
      double wv = WL[i];
      buf[wave][0][lane] = wv;
      buf[wave][1][lane] = wv * X[i];
      buf[wave][2][lane] = wv * Y[i];
      buf[wave][3][lane] = wv * Z[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int cnt = min(64, m - off);
    if (lane < 4) {
      const double* q = buf[wave][lane];
      for (int j = 0; j < cnt; j++) acc = acc + q[j];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (lane == 0) sw[c] = acc;
  if (lane == 1) sx[c] = acc;
  if (lane == 2) sy[c] = acc;
  if (lane == 3) sz[c] = acc;
}

// Division and empty-cluster compaction, utility.cpp:831-854 (swap-with-last,
// re-testing the slot that received the last cluster), by one block.
// Order in which the next centroid sums take their clusters: costliest first
// (cost as the last sums measured it, cost[map[c]] for cluster c after the
// compaction; map null: no compaction), by a 32-bucket histogram of log2 of
// the cost.  Any order gives the same sums; it only keeps the workgroups of
// the largest boxes from starting last.  Block-wide; ends with a barrier.
__device__ void order_by_cost(const int* __restrict__ cost, const int* __restrict__ map, int r,
                              int* __restrict__ order) {
  __shared__ int hist[32];
  if (threadIdx.x < 32) hist[threadIdx.x] = 0;
  __syncthreads();
  auto bucket = [&](int c) {
    const int v = cost[map ? map[c] : c];
    return v > 0 ? 31 - __clz(v) : 0;
  };
  for (int c = threadIdx.x; c < r; c += blockDim.x) atomicAdd(&hist[bucket(c)], 1);
  __syncthreads();
  if (threadIdx.x == 0) {  // exclusive offsets, bucket 31 first
    int o = 0;
    for (int k = 31; k >= 0; k--) {
      const int v = hist[k];
      hist[k] = o;
      o += v;
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < r; c += blockDim.x) order[atomicAdd(&hist[bucket(c)], 1)] = c;
  __syncthreads();
}

// Centroid update with the empty-cluster compaction (utility.cpp:824-856):
// new centroids from the sums, clusters that lost every point removed by
// swap-with-last (map: new index -> old).  Also resets the cluster boxes for
// the next assignment and, when order is given, orders the next sums.
__device__ void update_centroids(const double* __restrict__ sw, const double* __restrict__ sx,
                                 const double* __restrict__ sy, const double* __restrict__ sz,
                                 double* __restrict__ cx, double* __restrict__ cy,
                                 double* __restrict__ cz, int* __restrict__ map, int* scal,
                                 int* __restrict__ bbox, int rkmax, const int* __restrict__ cost,
                                 int* __restrict__ order) {
  __shared__ int nempty, rknew;
  const int rk = scal[KS_RK];
  if (bbox)  // the next assignment rebuilds the cluster boxes
    for (int e = threadIdx.x; e < 6 * rkmax; e += blockDim.x)
      bbox[e] = (e < 3 * rkmax) ? 0x7fffffff : -1;
  if (threadIdx.x == 0) nempty = 0;
  __syncthreads();
  int e = 0;
  for (int c = threadIdx.x; c < rk; c += blockDim.x) e += (sw[c] != 0.0) ? 0 : 1;
  if (e) atomicAdd(&nempty, e);
  __syncthreads();
  if (nempty == 0) {
    for (int c = threadIdx.x; c < rk; c += blockDim.x) {
      double d = sw[c];
      cx[c] = sx[c] / d;
      cy[c] = sy[c] / d;
      cz[c] = sz[c] / d;
    }
    if (threadIdx.x == 0) scal[KS_CHANGED] = 0;
    if (order) order_by_cost(cost, nullptr, rk, order);
    return;
  }
  if (threadIdx.x == 0) {
    for (int c = 0; c < rk; c++) map[c] = c;
    int r = rk;
    int i = 0;
    while (i < r) {
      if (sw[map[i]] != 0.0) {
        i++;
      } else {
        r--;
        map[i] = map[r];
      }
    }
    rknew = r;
  }
  __syncthreads();
  const int r = rknew;
  for (int c = threadIdx.x; c < r; c += blockDim.x) {
    int src = map[c];
    double d = sw[src];
    cx[c] = sx[src] / d;
    cy[c] = sy[src] / d;
    cz[c] = sz[src] / d;
  }
  if (threadIdx.x == 0) {
    scal[KS_RK] = r;
    scal[KS_CHANGED] = 0;
  }
  if (order) order_by_cost(cost, map, r, order);
}

template <bool BATCH>
__global__ __launch_bounds__(1024) void k_update(const double* __restrict__ sw,
                                                 const double* __restrict__ sx,
                                                 const double* __restrict__ sy,
                                                 const double* __restrict__ sz,
                                                 double* __restrict__ cx, double* __restrict__ cy,
                                                 double* __restrict__ cz, int* __restrict__ map,
                                                 int* scal, int* __restrict__ bbox, int rkmax,
                                                 const int* __restrict__ cost,
                                                 int* __restrict__ order, const KmJob* jobs) {
  if constexpr (BATCH) {
    const KmJob& J = jobs[blockIdx.y];
    sw = J.b.sw;
    sx = J.b.sx;
    sy = J.b.sy;
    sz = J.b.sz;
    cx = J.b.cx;
    cy = J.b.cy;
    cz = J.b.cz;
    map = J.b.map;
    scal = J.b.scal;
    bbox = J.b.bbox;
    rkmax = J.b.rkmax;
    order = nullptr;
  }
  if (scal[KS_ACTIVE] == 0) return;
  update_centroids(sw, sx, sy, sz, cx, cy, cz, map, scal, bbox, rkmax, cost, order);
}

// Loop control, utility.cpp:813 and :906-921, by one block: the changed-label
// count is the sum of the assignment's per-block counts.
__device__ void loop_control(int s, int max_iter, double maxF, int nblk,
                             const int* __restrict__ changed_blk, int* trace, int* scal) {
  int v = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) v += changed_blk[b];
  int f;
  block_excl_scan(v, &f);
  if (threadIdx.x == 0) {
    trace[s] = f;
    scal[KS_ITERS] = s + 1;
    if (!((double)f > maxF && s + 1 < max_iter)) scal[KS_ACTIVE] = 0;
  }
}

template <bool BATCH>
__global__ __launch_bounds__(1024) void k_control(int s, int max_iter, double maxF, int nblk,
                                                  const int* __restrict__ changed_blk, int* trace,
                                                  int* scal, const KmJob* jobs) {
  if constexpr (BATCH) {
    const KmJob& J = jobs[blockIdx.y];
    max_iter = J.max_iter;
    maxF = J.maxF;
    nblk = J.b.nblk_pts;
    changed_blk = J.b.changed_blk;
    trace = J.b.trace;
    scal = J.b.scal;
  }
  if (scal[KS_ACTIVE] == 0) return;
  loop_control(s, max_iter, maxF, nblk, changed_blk, trace, scal);
}

// The loop control of iteration s-1 and, if the loop goes on, the centroid
// update of iteration s, in one launch queued after the sums of iteration s.
// Those sums read only the labels of iteration s-1, which the control does not
// touch; when the control ends the loop, the sums are simply not used.
// Batched: one block per instance (blockIdx.y), each with its own iteration
// limit and buffers.
template <bool BATCH>
__global__ __launch_bounds__(1024) void k_control_update(
    int s, int max_iter, double maxF, int nblk, const int* __restrict__ changed_blk, int* trace,
    const double* __restrict__ sw, const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, double* __restrict__ cx, double* __restrict__ cy,
    double* __restrict__ cz, int* __restrict__ map, int* scal, int* __restrict__ bbox,
    int rkmax, const int* __restrict__ cost, int* __restrict__ order, const KmJob* jobs) {
  if constexpr (BATCH) {
    const KmJob& J = jobs[blockIdx.y];
    max_iter = J.max_iter;
    maxF = J.maxF;
    nblk = J.b.nblk_pts;
    changed_blk = J.b.changed_blk;
    trace = J.b.trace;
    sw = J.b.sw;
    sx = J.b.sx;
    sy = J.b.sy;
    sz = J.b.sz;
    cx = J.b.cx;
    cy = J.b.cy;
    cz = J.b.cz;
    map = J.b.map;
    scal = J.b.scal;
    bbox = J.b.bbox;
    rkmax = J.b.rkmax;
    cost = nullptr;
    order = nullptr;
  }
  // Every load that does not depend on another goes out first: the flag and
  // rk, this thread's changed-label counts, and the sums and costs of its
  // clusters (c = t + 1024 u; indices below rkmax are in the buffers whatever
  // rk is), so the kernel waits for memory about once instead of once per step.
  constexpr int PF = 4;
  const int t = threadIdx.x;
  const int active = scal[KS_ACTIVE];
  const int rk = scal[KS_RK];
  int v = 0;
  for (int b = t; b < nblk; b += blockDim.x) v += changed_blk[b];
  const bool pre = rkmax <= PF * (int)blockDim.x;
  double pw[PF], px[PF], py[PF], pz[PF];
  int pc[PF];
#pragma unroll
  for (int u = 0; u < PF; u++) {
    const int c = t + u * (int)blockDim.x;
    const bool ok = pre && c < rkmax;
    pw[u] = ok ? sw[c] : 1.0;
    px[u] = ok ? sx[c] : 0.0;
    py[u] = ok ? sy[c] : 0.0;
    pz[u] = ok ? sz[c] : 0.0;
    pc[u] = (ok && cost) ? cost[c] : 0;
  }
  if (active == 0) return;
  // Loop control of iteration s-1 (utility.cpp:813, :906-921).
  int f;
  block_excl_scan(v, &f);
  const bool go = (double)f > maxF && s < max_iter;
  if (t == 0) {
    trace[s - 1] = f;
    scal[KS_ITERS] = s;
    if (!go) scal[KS_ACTIVE] = 0;
  }
  if (!go) return;
  if (!pre) {
    update_centroids(sw, sx, sy, sz, cx, cy, cz, map, scal, bbox, rkmax, cost, order);
    return;
  }
  // A cluster that lost every point takes the compaction path, which reads the
  // sums again; otherwise the update runs from the registers.
  int empty = 0;
#pragma unroll
  for (int u = 0; u < PF; u++) {
    const int c = t + u * (int)blockDim.x;
    if (c < rk && pw[u] == 0.0) empty = 1;
  }
  if (__syncthreads_or(empty)) {
    update_centroids(sw, sx, sy, sz, cx, cy, cz, map, scal, bbox, rkmax, cost, order);
    return;
  }
  if (bbox)  // the next assignment rebuilds the cluster boxes
    for (int e = t; e < 6 * rkmax; e += blockDim.x) bbox[e] = (e < 3 * rkmax) ? 0x7fffffff : -1;
#pragma unroll
  for (int u = 0; u < PF; u++) {
    const int c = t + u * (int)blockDim.x;
    if (c < rk) {
      cx[c] = px[u] / pw[u];
      cy[c] = py[u] / pw[u];
      cz[c] = pz[u] / pw[u];
    }
  }
  if (t == 0) scal[KS_CHANGED] = 0;
  if (order) {  // order_by_cost with the costs in registers
    __shared__ int hist[32];
    if (t < 32) hist[t] = 0;
    __syncthreads();
    int bk[PF];
#pragma unroll
    for (int u = 0; u < PF; u++) {
      bk[u] = pc[u] > 0 ? 31 - __clz(pc[u]) : 0;
      if (t + u * (int)blockDim.x < rk) atomicAdd(&hist[bk[u]], 1);
    }
    __syncthreads();
    if (t == 0) {  // exclusive offsets, bucket 31 first
      int o = 0;
      for (int k = 31; k >= 0; k--) {
        const int hv = hist[k];
        hist[k] = o;
        o += hv;
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PF; u++) {
      const int c = t + u * (int)blockDim.x;
      if (c < rk) order[atomicAdd(&hist[bk[u]], 1)] = c;
    }
  }
}

// findMin(Dim=0), utility.cpp:926-945: per centroid, the first point at the
// minimum distance.  Lanes own centroids; a block walks one chunk of points.
__global__ __launch_bounds__(256) void k_final_partial(int n, int chunk,
                                                       const double* __restrict__ X,
                                                       const double* __restrict__ Y,
                                                       const double* __restrict__ Z,
                                                       const double* __restrict__ cx,
                                                       const double* __restrict__ cy,
                                                       const double* __restrict__ cz,
                                                       double* __restrict__ pbest,
                                                       int* __restrict__ pidx, const int* scal) {
  const int rk = scal[KS_RK];
  if ((int)(blockIdx.y * 256) >= rk) return;
  const int c = blockIdx.y * 256 + threadIdx.x;
  const bool ok = c < rk;
  const int cc = ok ? c : 0;
  const double qx = cx[cc], qy = cy[cc], qz = cz[cc];
  const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
  double best = dist2(X[i0], Y[i0], Z[i0], qx, qy, qz);
  int bi = i0;
  for (int i = i0 + 1; i < i1; i++) {
    double d = dist2(X[i], Y[i], Z[i], qx, qy, qz);
    if (d < best) { best = d; bi = i; }
  }
  if (ok) {
    pbest[(long long)blockIdx.x * rk + c] = best;
    pidx[(long long)blockIdx.x * rk + c] = bi;
  }
}

__global__ __launch_bounds__(256) void k_final_reduce(int nchunk, const double* __restrict__ pbest,
                                                      const int* __restrict__ pidx,
                                                      const int* __restrict__ gidx,
                                                      int* __restrict__ imin, const int* scal) {
  const int rk = scal[KS_RK];
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= rk) return;
  double best = pbest[c];
  int bi = pidx[c];
  for (int b = 1; b < nchunk; b++) {
    double v = pbest[(long long)b * rk + c];
    if (v < best) {
      best = v;
      bi = pidx[(long long)b * rk + c];
    }
  }
  imin[c] = gidx[bi];
}

// unique(), utility.cpp:617-629: bitonic sort in LDS, then drop repeats.
template <bool BATCH>
__global__ __launch_bounds__(1024) void k_unique(const int* __restrict__ imin,
                                                 int* __restrict__ centers, int* scal,
                                                 const KmJob* jobs) {
  extern __shared__ int s[];
  if constexpr (BATCH) {
    const KmJob& J = jobs[blockIdx.y];
    imin = J.b.imin;
    centers = J.b.centers;
    scal = J.b.scal;
  }
  const int rk = scal[KS_RK];
  int P = 1;
  while (P < rk) P <<= 1;
  for (int i = threadIdx.x; i < P; i += blockDim.x) s[i] = (i < rk) ? imin[i] : 0x7fffffff;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        int l = i ^ j;
        if (l > i) {
          int a = s[i], b = s[l];
          bool up = (i & k) == 0;
          if ((a > b) == up) {
            s[i] = b;
            s[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int off = 0; off < rk; off += blockDim.x) {
    int i = off + threadIdx.x;
    int keep = (i < rk && (i == 0 || s[i] != s[i - 1])) ? 1 : 0;
    int tot;
    int ex = block_excl_scan(keep, &tot);
    if (keep) centers[carry + ex] = s[i];
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) scal[KS_RKU] = carry;
}

template <bool BATCH>
__global__ __launch_bounds__(256) void k_mark_centers(const int* __restrict__ centers,
                                                      uint8_t* __restrict__ flags,
                                                      int* __restrict__ piv, const int* scal,
                                                      const KmJob* jobs) {
  if constexpr (BATCH) {
    const KmJob& J = jobs[blockIdx.y];
    centers = J.b.centers;
    flags = J.b.flags;
    piv = J.piv;
    scal = J.b.scal;
  }
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= scal[KS_RKU]) return;
  int g = centers[k];
  flags[g] = 1;
  piv[k] = g;
}

__global__ void k_fill_u8(uint8_t* p, long long n, uint8_t v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_i32(int* p, long long n, int v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

// Ordered compaction of the indices whose flag equals `want` into out[base...].
// Returns the count through the device slot `count_dev` (bcount[nblk]).
// Batched helpers (blockIdx.y = instance).
__global__ void k_km_batch_init(const KmJob* __restrict__ jobs) {
  const KmJob& J = jobs[blockIdx.y];
  if (threadIdx.x < KS_COUNT) J.b.scal[threadIdx.x] = 0;
}

// flag = n before the first test (utility.cpp:765, :813).
// Also keeps npt (the threshold count) in scal[KS_NPT] before the final
// compaction reuses bcount.
__global__ void k_km_batch_active(const KmJob* __restrict__ jobs) {
  const KmJob& J = jobs[blockIdx.y];
  if (threadIdx.x == 0) {
    J.b.scal[KS_ACTIVE] = ((double)J.b.n > J.maxF && 0 < J.max_iter) ? 1 : 0;
    J.b.scal[KS_NPT] = J.b.bcount[J.b.nblk_flag];
  }
}

__global__ void k_km_batch_clear_flags(const KmJob* __restrict__ jobs) {
  const KmJob& J = jobs[blockIdx.y];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < J.b.n;
       i += (long long)gridDim.x * blockDim.x)
    J.b.flags[i] = 0;
}

// One assignment of a single instance: bricks of kBrickPZ planes per thread in
// grid mode, runs of 256 points otherwise; sl / sc: super-brick lists or null.
template <int CAP>
void launch_assign(KmeansBufs* b, int count_changes, const int* sl, const int* sc, hipStream_t st) {
  if (b->grid_mode)
    hipLaunchKernelGGL((k_assign_culled<false, CAP, kBrickPZ>), dim3(b->nblk_pts), dim3(256), 0, st,
                       (const int*)nullptr, b->X, b->Y, b->Z, b->cx, b->cy, b->cz, b->label,
                       b->scal, count_changes, b->bbox, b->rkmax, 1, b->ngrid[0], b->ngrid[1],
                       b->ngrid[2], b->dx[0], b->dx[1], b->dx[2], b->changed_blk, sl, b->slxyz,
                       sc, b->nsb, (const KmJob*)nullptr);
  else
    hipLaunchKernelGGL((k_assign_culled<false, CAP, 1>), dim3(b->nblk_pts), dim3(256), 0, st,
                       b->porder, b->X, b->Y, b->Z, b->cx, b->cy, b->cz, b->label, b->scal,
                       count_changes, b->bbox, b->rkmax, 0, b->ngrid[0], b->ngrid[1],
                       b->ngrid[2], b->dx[0], b->dx[1], b->dx[2], b->changed_blk,
                       (const int*)nullptr, (const double*)nullptr, (const int*)nullptr, 0,
                       (const KmJob*)nullptr);
}

int compact_flags(KmeansBufs* b, int want, int base, int* out, hipStream_t st,
                  const int* base_dev = nullptr) {
  const int nblk = b->nblk_flag;
  hipLaunchKernelGGL(k_flag_count<false>, dim3(nblk), dim3(256), 0, st, b->flags, b->n, want, b->bcount, (const KmJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_flag_count");
  hipLaunchKernelGGL(k_scan_single<false>, dim3(1), dim3(1024), 0, st, b->bcount, nblk, (int*)nullptr,
                     (const KmJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_scan_single");
  if (out) {
    hipLaunchKernelGGL(k_flag_scatter<false>, dim3(nblk), dim3(256), 0, st, b->flags, b->n, want,
                       b->bcount, base, out, base_dev, (const KmJob*)nullptr);
    ISDF_HIP_LAUNCHED("k_flag_scatter");
  }
  return 0;
}

}  // namespace

size_t kmeans_bytes(int n, const int ngrid[3], int rkmax, int max_iter) {
  KmeansBufs b;
  kmeans_carve(&b, nullptr, n, ngrid, rkmax, max_iter);
  return (size_t)(uintptr_t)b.scal + 64;
}

void kmeans_carve(KmeansBufs* b, void* basep, int n, const int ngrid[3], int rkmax, int max_iter) {
  uintptr_t p = (uintptr_t)basep;
  auto take = [&](size_t bytes) {
    uintptr_t r = p;
    p += (bytes + 255) & ~(size_t)255;
    return r;
  };
  b->n = n;
  b->rkmax = rkmax;
  b->chunk_sort = kSortChunk;
  b->nchunk_sort = (int)cdiv(n, kSortChunk);
  int tiles = (int)cdiv(rkmax, 256);
  int want = std::max(1, 2048 / tiles);
  int chunk = std::max(256, (int)cdiv(n, want));
  b->chunk_final = chunk;
  b->nchunk_final = (int)cdiv(n, chunk);
  b->nblk_flag = (int)cdiv(n, kFlagTile);
  b->w = (double*)take(sizeof(double) * n);
  b->flags = (uint8_t*)take((size_t)n);
  b->index = (int*)take(sizeof(int) * n);
  b->gidx = (int*)take(sizeof(int) * n);
  b->X = (double*)take(sizeof(double) * n);
  b->Y = (double*)take(sizeof(double) * n);
  b->Z = (double*)take(sizeof(double) * n);
  b->WL = (double*)take(sizeof(double) * n);
  b->label = (int*)take(sizeof(int) * n);
  b->sorted = (int*)take(sizeof(int) * n);
  b->bcount = (int*)take(sizeof(int) * (b->nblk_flag + 1));
  b->hist = (int*)take(sizeof(int) * (size_t)b->nchunk_sort * rkmax);
  b->cnt_pts = (int*)take(sizeof(int) * rkmax);
  b->base = (int*)take(sizeof(int) * (rkmax + 1));
  b->map = (int*)take(sizeof(int) * rkmax);
  b->cx = (double*)take(sizeof(double) * rkmax);
  b->cy = (double*)take(sizeof(double) * rkmax);
  b->cz = (double*)take(sizeof(double) * rkmax);
  b->sw = (double*)take(sizeof(double) * rkmax);
  b->sx = (double*)take(sizeof(double) * rkmax);
  b->sy = (double*)take(sizeof(double) * rkmax);
  b->sz = (double*)take(sizeof(double) * rkmax);
  b->pbest = (double*)take(sizeof(double) * (size_t)b->nchunk_final * rkmax);
  b->pidx = (int*)take(sizeof(int) * (size_t)b->nchunk_final * rkmax);
  b->imin = (int*)take(sizeof(int) * rkmax);
  b->centers = (int*)take(sizeof(int) * rkmax);
  b->trace = (int*)take(sizeof(int) * std::max(1, max_iter));
  b->nblk_pts = std::max((int)cdiv(n, 256), 1);
  {
    // point-block table: the larger of the brick tiling and 256-point runs
    const long long bricks =
        (long long)((ngrid[0] + 7) / 8) * ((ngrid[1] + 7) / 8) * ((ngrid[2] + 3) / 4);
    const long long blocks = std::max<long long>(bricks, (long long)cdiv(n, 256));
    b->porder = (int*)take(sizeof(int) * (size_t)(blocks * 256));
    b->changed_blk = (int*)take(sizeof(int) * (size_t)blocks);
  }
  b->bbox = (int*)take(sizeof(int) * 6 * (size_t)rkmax);
  b->cost = (int*)take(sizeof(int) * (size_t)rkmax);
  b->order = (int*)take(sizeof(int) * (size_t)rkmax);
  b->nsb = ((ngrid[0] + 15) / 16) * ((ngrid[1] + 15) / 16) * ((ngrid[2] + 7) / 8);
  // Super-brick candidate lists, used from rk = kSuperMinRk on.
  const bool super = rkmax >= kSuperMinRk;
  b->slist = super ? (int*)take(sizeof(int) * (size_t)b->nsb * kSuperCap) : nullptr;
  b->slxyz = super ? (double*)take(sizeof(double) * 3 * (size_t)b->nsb * kSuperCap) : nullptr;
  b->scount = (int*)take(sizeof(int) * (size_t)b->nsb);
  b->scal = (int*)take(sizeof(int) * KS_COUNT);
}

int launch_weights(const double* psi, long long ld, int ng, int s1lo, int s1hi, int s2lo,
                   int s2hi, double* w, hipStream_t st) {
  int lo = std::min(s1lo, s2lo), hi = std::max(s1hi, s2hi);
  hipLaunchKernelGGL(k_weights<false>, dim3(cdiv(ng, 256)), dim3(256), 0, st, psi, ld, ng, lo, hi, s1lo,
                     s1hi, s2lo, s2hi, w, (const KmJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_weights");
  return 0;
}

int kmeans_prepare(KmeansBufs* b, const int ngrid[3], const double length[3], double df_tol,
                   int* npt_host, hipStream_t st) {
  (void)ngrid;
  (void)length;
  const int n = b->n;
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, b->scal, (long long)KS_COUNT, 0);
  ISDF_HIP_LAUNCHED("k_fill_i32");
  unsigned long long* maxbits = (unsigned long long*)(b->scal + KS_MAXW);
  hipLaunchKernelGGL(k_maxw<false>, dim3(std::min(cdiv(n, 256), 256u)), dim3(256), 0, st, b->w, n,
                     maxbits, (const KmJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_maxw");
  hipLaunchKernelGGL(k_thresh_count<false>, dim3(b->nblk_flag), dim3(256), 0, st, b->w, n,
                     (const unsigned long long*)maxbits, df_tol, b->flags, b->bcount,
                     (const KmJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_thresh_count");
  hipLaunchKernelGGL(k_scan_single<false>, dim3(1), dim3(1024), 0, st, b->bcount, b->nblk_flag,
                     b->scal + KS_NPT, (const KmJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_scan_single");
  // Without npt_host the caller assumes every grid point is a K-means point and
  // checks npt (scal[KS_NPT]) afterwards.
  if (!npt_host) return 0;
  int npt = 0;
  ISDF_HIP_TRY(hipMemcpyAsync(&npt, b->bcount + b->nblk_flag, sizeof(int), hipMemcpyDeviceToHost,
                              st));
  ISDF_HIP_TRY(hipStreamSynchronize(st));
  *npt_host = npt;
  if (npt < n) {
    hipLaunchKernelGGL(k_flag_scatter<false>, dim3(b->nblk_flag), dim3(256), 0, st, b->flags, n, 1,
                       b->bcount, 0, b->index, (const int*)nullptr, (const KmJob*)nullptr);
    ISDF_HIP_LAUNCHED("k_flag_scatter");
  }
  return 0;
}

int kmeans_grid_bricks(const int ngrid[3]) {
  const int nbx = (ngrid[0] + 7) / 8, nby = (ngrid[1] + 7) / 8;
  const int nbz = (ngrid[2] + 4 * kBrickPZ - 1) / (4 * kBrickPZ);
  return nbx * nby * nbz;
}

void kmeans_geometry(KmeansBufs* b, const int ngrid[3], const double length[3], int npt) {
  for (int d = 0; d < 3; d++) {
    b->ngrid[d] = ngrid[d];
    b->dx[d] = length[d] / ngrid[d];
  }
  b->grid_mode = (npt == b->n);
  if (b->grid_mode) {
    b->nblk_pts = kmeans_grid_bricks(ngrid);
  } else {
    b->nblk_pts = (int)cdiv(b->n, 256);
  }
}

int kmeans_setup_points(KmeansBufs* b, const int ngrid[3], const double length[3], int npt,
                        int rk, hipStream_t st) {
  const int n = b->n;
  double dx0 = length[0] / ngrid[0], dx1 = length[1] / ngrid[1], dx2 = length[2] / ngrid[2];
  kmeans_geometry(b, ngrid, length, npt);
  hipLaunchKernelGGL(k_setup_points<false>, dim3(cdiv(n, 256)), dim3(256), 0, st, n, npt, b->index, b->w,
                     ngrid[0], ngrid[1], dx0, dx1, dx2, b->gidx, b->X, b->Y, b->Z, b->WL, (const KmJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_setup_points");
  hipLaunchKernelGGL(k_init_centroids<false>, dim3(cdiv(rk, 256)), dim3(256), 0, st, rk, n, npt, b->index,
                     ngrid[0], ngrid[1], dx0, dx1, dx2, b->cx, b->cy, b->cz, b->scal, (const KmJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_init_centroids");
  if (!b->grid_mode) {  // grid mode: the assignment numbers the bricks itself
    hipLaunchKernelGGL(k_build_blocks<false>, dim3(b->nblk_pts), dim3(256), 0, st, n, 0, ngrid[0],
                       ngrid[1], ngrid[2], b->porder, (const KmJob*)nullptr);
    ISDF_HIP_LAUNCHED("k_build_blocks");
  }
  hipLaunchKernelGGL(k_reset_boxes<false>, dim3(cdiv(6 * (long long)b->rkmax, 256)), dim3(256), 0, st,
                     b->bbox, b->rkmax, (const KmJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_reset_boxes");
  launch_assign<kCandCapWide>(b, 0, nullptr, nullptr, st);
  ISDF_HIP_LAUNCHED("k_assign_culled");
  return 0;
}

int kmeans_iterate(KmeansBufs* b, int rk, int max_iter, double kmeans_tol, int* iters_host,
                   hipStream_t st) {
  const int n = b->n;
  const double maxF = kmeans_tol * n;
  // flag = n before the first test (utility.cpp:765, :813).
  const int active0 = ((double)n > maxF && 0 < max_iter) ? 1 : 0;
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, b->scal + KS_ACTIVE, 1LL, active0);
  ISDF_HIP_LAUNCHED("k_fill_i32");
  const bool poll = iters_host != nullptr;
  if (poll) *iters_host = 0;
  if (!active0) return 0;
  const size_t lds_rk = sizeof(int) * (size_t)rk;
  // Polled: the host reads the active flag every few iterations and stops
  // queueing; unpolled (graph capture): every iteration is queued and the
  // kernels return at once after convergence.
  const int check_every = 4;
  // Per iteration s: the sums, then the loop control of iteration s-1 with the
  // centroid update of iteration s (one launch), then the assignment.
  int last = 0;
  // Grid mode: each centroid-sum launch records its clusters' costs, and the
  // update after it orders the next launch's clusters by them.
  int* const cost = b->grid_mode ? b->cost : nullptr;
  for (int s = 0; s < max_iter; s++) {
    const int* order = (b->grid_mode && s > 0) ? b->order : nullptr;
    if (b->grid_mode && s < kWideSumsIters) {
      hipLaunchKernelGGL((k_centroid_sums_box<false, 15, 4>), dim3(rk), dim3(1024), 0, st,
                         b->ngrid[0], b->ngrid[1], b->label, b->WL, b->X, b->Y, b->Z, b->bbox,
                         b->rkmax, b->sw, b->sx, b->sy, b->sz, b->scal, cost, order,
                         (const KmJob*)nullptr);
    } else if (b->grid_mode) {
      hipLaunchKernelGGL((k_centroid_sums_box<false, 4, 2>), dim3(rk), dim3(320), 0, st, b->ngrid[0],
                         b->ngrid[1], b->label, b->WL, b->X, b->Y, b->Z, b->bbox, b->rkmax, b->sw,
                         b->sx, b->sy, b->sz, b->scal, cost, order, (const KmJob*)nullptr);
    } else {
      hipLaunchKernelGGL(k_sort_hist, dim3(b->nchunk_sort), dim3(256), lds_rk, st, n, b->label,
                         b->hist, b->scal);
      hipLaunchKernelGGL(k_sort_scan_chunks, dim3(cdiv(rk, 256)), dim3(256), 0, st,
                         b->nchunk_sort, b->hist, b->cnt_pts, b->scal);
      hipLaunchKernelGGL(k_sort_scan_clusters, dim3(1), dim3(1024), 0, st, b->cnt_pts, b->base,
                         b->scal);
      hipLaunchKernelGGL(k_sort_scatter, dim3(b->nchunk_sort), dim3(64), lds_rk, st, n, b->label,
                         b->hist, b->base, b->sorted, b->scal);
      hipLaunchKernelGGL(k_centroid_sums, dim3(cdiv(rk, 4)), dim3(256), 0, st, b->sorted,
                         b->base, b->cnt_pts, b->WL, b->X, b->Y, b->Z, b->sw, b->sx, b->sy, b->sz,
                         b->scal);
    }
    if (s == 0)
      hipLaunchKernelGGL(k_update<false>, dim3(1), dim3(1024), 0, st, b->sw, b->sx, b->sy, b->sz,
                         b->cx, b->cy, b->cz, b->map, b->scal,
                         b->grid_mode ? b->bbox : (int*)nullptr, b->rkmax, cost,
                         cost ? b->order : (int*)nullptr, (const KmJob*)nullptr);
    else
      hipLaunchKernelGGL(k_control_update<false>, dim3(1), dim3(1024), 0, st, s, max_iter, maxF,
                         b->nblk_pts, b->changed_blk, b->trace, b->sw, b->sx, b->sy, b->sz, b->cx,
                         b->cy, b->cz, b->map, b->scal, b->grid_mode ? b->bbox : (int*)nullptr,
                         b->rkmax, cost, cost ? b->order : (int*)nullptr, (const KmJob*)nullptr);
    last = s;
    // Grid mode: the super-brick candidate lists first (the bricks cull from
    // them).  The first iterations keep the wide candidate list; their
    // centroids still crowd the corner of the initial ones.
    const int* sl = nullptr;
    const int* sc = nullptr;
    if (b->grid_mode && s >= 3 && rk >= kSuperMinRk) {
      hipLaunchKernelGGL(k_cull_super, dim3(b->nsb), dim3(1024), 0, st, b->cx, b->cy, b->cz,
                         b->ngrid[0], b->ngrid[1], b->ngrid[2], b->dx[0], b->dx[1], b->dx[2],
                         b->slist, b->slxyz, b->scount, b->scal);
      sl = b->slist;
      sc = b->scount;
    }
    if (s == 0)
      launch_assign<kCandCapWide>(b, 1, sl, sc, st);
    else
      launch_assign<kCandCap>(b, 1, sl, sc, st);
    ISDF_HIP_LAUNCHED("k-means iteration");
    if (poll && (s + 1) % check_every == 0 && s + 1 < max_iter) {
      int act = 1;
      ISDF_HIP_TRY(hipMemcpyAsync(&act, b->scal + KS_ACTIVE, sizeof(int), hipMemcpyDeviceToHost,
                                  st));
      ISDF_HIP_TRY(hipStreamSynchronize(st));
      if (!act) break;
    }
  }
  // The loop control of the last queued assignment.
  hipLaunchKernelGGL(k_control<false>, dim3(1), dim3(1024), 0, st, last, max_iter, maxF,
                     b->nblk_pts, b->changed_blk, b->trace, b->scal, (const KmJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_control");
  if (!poll) return 0;
  int it = 0;
  ISDF_HIP_TRY(hipMemcpyAsync(&it, b->scal + KS_ITERS, sizeof(int), hipMemcpyDeviceToHost, st));
  ISDF_HIP_TRY(hipStreamSynchronize(st));
  *iters_host = it;
  return 0;
}

int kmeans_finish(KmeansBufs* b, int* piv_dev, hipStream_t st, bool nosync) {
  const int n = b->n;
  const int rkmax = b->rkmax;
  if (b->grid_mode) {
    hipLaunchKernelGGL(k_final_window<false>, dim3(cdiv(rkmax, 4)), dim3(256), 0, st, b->ngrid[0],
                       b->ngrid[1], b->ngrid[2], b->dx[0], b->dx[1], b->dx[2], b->X, b->Y, b->Z,
                       b->cx, b->cy, b->cz, b->imin, b->scal, (const KmJob*)nullptr);
    ISDF_HIP_LAUNCHED("k_final_window");
  } else {
    dim3 g(b->nchunk_final, cdiv(rkmax, 256));
    hipLaunchKernelGGL(k_final_partial, g, dim3(256), 0, st, n, b->chunk_final, b->X, b->Y, b->Z,
                       b->cx, b->cy, b->cz, b->pbest, b->pidx, b->scal);
    ISDF_HIP_LAUNCHED("k_final_partial");
    hipLaunchKernelGGL(k_final_reduce, dim3(cdiv(rkmax, 256)), dim3(256), 0, st, b->nchunk_final,
                       b->pbest, b->pidx, b->gidx, b->imin, b->scal);
    ISDF_HIP_LAUNCHED("k_final_reduce");
  }
  int P = 1;
  while (P < rkmax) P <<= 1;
  hipLaunchKernelGGL(k_unique<false>, dim3(1), dim3(1024), sizeof(int) * (size_t)P, st, b->imin,
                     b->centers, b->scal, (const KmJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_unique");
  // piv = [sorted unique centers] ++ [all other grid indices ascending] (utility.cpp:973-983).
  hipLaunchKernelGGL(k_fill_u8, dim3(std::min(cdiv(n, 256), 4096u)), dim3(256), 0, st, b->flags,
                     (long long)n, (uint8_t)0);
  hipLaunchKernelGGL(k_mark_centers<false>, dim3(cdiv(rkmax, 256)), dim3(256), 0, st, b->centers,
                     b->flags, piv_dev, b->scal, (const KmJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_mark_centers");
  if (nosync) return compact_flags(b, 0, 0, piv_dev, st, b->scal + KS_RKU);
  int rku = 0;
  ISDF_HIP_TRY(hipMemcpyAsync(&rku, b->scal + KS_RKU, sizeof(int), hipMemcpyDeviceToHost, st));
  ISDF_HIP_TRY(hipStreamSynchronize(st));
  return compact_flags(b, 0, rku, piv_dev, st);
}

int kmeans_batched(const KmJob* d_jobs, const KmJob* h_jobs, int nb, const int* d_bricks,
                   int nbricks, const int* d_clusters, int nclusters, hipStream_t st) {
  int maxn = 0, maxrk = 0, maxflag = 0, maxit = 0;
  long long bricks = 0, clusters = 0;
  for (int i = 0; i < nb; i++) {
    const KmeansBufs& b = h_jobs[i].b;
    if (!b.grid_mode) return set_error(1, "batched K-means needs every grid point as a point");
    maxn = std::max(maxn, b.n);
    maxrk = std::max(maxrk, b.rkmax);
    maxflag = std::max(maxflag, b.nblk_flag);
    bricks += b.nblk_pts;
    clusters += b.rkmax;
    maxit = std::max(maxit, h_jobs[i].max_iter);
  }
  if (bricks != nbricks || !d_bricks || clusters != nclusters || !d_clusters)
    return set_error(1, "batched K-means: the block lists do not match the instances");
  const KmJob* J = d_jobs;
  const unsigned gn = cdiv(maxn, 256);
  // Weights (getpoints_Kmeans_MPI.cpp:33-45).
  hipLaunchKernelGGL(k_weights<true>, dim3(gn, nb), dim3(256), 0, st, (const double*)nullptr, 0LL, 0, 0,
                     0, 0, 0, 0, 0, (double*)nullptr, J);
  ISDF_HIP_LAUNCHED("k_weights (batched)");
  // Threshold and npt (utility.cpp:646-662); npt stays in bcount[nblk_flag].
  hipLaunchKernelGGL(k_km_batch_init, dim3(1, nb), dim3(64), 0, st, J);
  hipLaunchKernelGGL(k_maxw<true>, dim3(std::min(gn, 256u), nb), dim3(256), 0, st,
                     (const double*)nullptr, 0, (unsigned long long*)nullptr, J);
  hipLaunchKernelGGL(k_thresh_count<true>, dim3(maxflag, nb), dim3(256), 0, st,
                     (const double*)nullptr, 0, (const unsigned long long*)nullptr, 1e-8,
                     (uint8_t*)nullptr, (int*)nullptr, J);
  hipLaunchKernelGGL(k_scan_single<true>, dim3(1, nb), dim3(1024), 0, st, (int*)nullptr, 0,
                     (int*)nullptr, J);
  ISDF_HIP_LAUNCHED("threshold (batched)");
  // Points, initial centroids, brick order, first assignment (utility.cpp:682-812).
  hipLaunchKernelGGL(k_setup_points<true>, dim3(gn, nb), dim3(256), 0, st, 0, 0, (const int*)nullptr,
                     (const double*)nullptr, 0, 0, 0.0, 0.0, 0.0, (int*)nullptr, (double*)nullptr,
                     (double*)nullptr, (double*)nullptr, (double*)nullptr, J);
  hipLaunchKernelGGL(k_init_centroids<true>, dim3(cdiv(maxrk, 256), nb), dim3(256), 0, st, 0, 0, 0,
                     (const int*)nullptr, 0, 0, 0.0, 0.0, 0.0, (double*)nullptr, (double*)nullptr,
                     (double*)nullptr, (int*)nullptr, J);
  hipLaunchKernelGGL(k_reset_boxes<true>, dim3(cdiv(6LL * maxrk, 256), nb), dim3(256), 0, st,
                     (int*)nullptr, 0, J);
  hipLaunchKernelGGL((k_assign_culled<true, kCandCapWide, kBrickPZ>), dim3(nbricks), dim3(256), 0, st,
                     d_bricks,
                     (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                     (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                     (int*)nullptr, (int*)nullptr, 0, (int*)nullptr, 0, 1, 0, 0, 0, 0.0, 0.0, 0.0, (int*)nullptr,
                     (const int*)nullptr, (const double*)nullptr, (const int*)nullptr, 0, J);
  hipLaunchKernelGGL(k_km_batch_active, dim3(1, nb), dim3(64), 0, st, J);
  ISDF_HIP_LAUNCHED("k-means setup (batched)");
  // Lloyd iterations (utility.cpp:812-922); converged instances return at once.
  // As in kmeans_iterate: the sums of iteration s, then the loop control of
  // iteration s-1 with the update of iteration s in one launch, then the
  // assignment; the control of the last iteration after the loop.
  for (int s = 0; s < maxit; s++) {
    hipLaunchKernelGGL((k_centroid_sums_box<true, 4, 2>), dim3(nclusters), dim3(320), 0, st, 0, 0,
                       (const int*)nullptr, (const double*)nullptr, (const double*)nullptr,
                       (const double*)nullptr, (const double*)nullptr, (const int*)nullptr, 0,
                       (double*)nullptr, (double*)nullptr, (double*)nullptr, (double*)nullptr,
                       (const int*)nullptr, (int*)nullptr, d_clusters, J);
    if (s == 0)
      hipLaunchKernelGGL(k_update<true>, dim3(1, nb), dim3(1024), 0, st, (const double*)nullptr,
                         (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                         (double*)nullptr, (double*)nullptr, (double*)nullptr, (int*)nullptr,
                         (int*)nullptr, (int*)nullptr, 0, (const int*)nullptr, (int*)nullptr, J);
    else
      hipLaunchKernelGGL(k_control_update<true>, dim3(1, nb), dim3(1024), 0, st, s, 0, 0.0, 0,
                         (const int*)nullptr, (int*)nullptr, (const double*)nullptr,
                         (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                         (double*)nullptr, (double*)nullptr, (double*)nullptr, (int*)nullptr,
                         (int*)nullptr, (int*)nullptr, 0, (const int*)nullptr, (int*)nullptr, J);
    if (s == 0)
      hipLaunchKernelGGL((k_assign_culled<true, kCandCapWide, kBrickPZ>), dim3(nbricks), dim3(256), 0, st, d_bricks,
                         (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                         (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                         (int*)nullptr, (int*)nullptr, 1, (int*)nullptr, 0, 1, 0, 0, 0, 0.0, 0.0, 0.0, (int*)nullptr,
                         (const int*)nullptr, (const double*)nullptr, (const int*)nullptr, 0, J);
    else
      hipLaunchKernelGGL((k_assign_culled<true, kCandCap, kBrickPZ>), dim3(nbricks), dim3(256), 0, st, d_bricks,
                         (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                         (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                         (int*)nullptr, (int*)nullptr, 1, (int*)nullptr, 0, 1, 0, 0, 0, 0.0, 0.0, 0.0, (int*)nullptr,
                         (const int*)nullptr, (const double*)nullptr, (const int*)nullptr, 0, J);
    ISDF_HIP_LAUNCHED("k-means iteration (batched)");
  }
  if (maxit > 0)
    hipLaunchKernelGGL(k_control<true>, dim3(1, nb), dim3(1024), 0, st, maxit - 1, 0, 0.0, 0,
                       (const int*)nullptr, (int*)nullptr, (int*)nullptr, J);
  // Final nearest points, unique(), piv (utility.cpp:926-983).
  hipLaunchKernelGGL(k_final_window<true>, dim3(cdiv(maxrk, 4), nb), dim3(256), 0, st, 0, 0, 0, 0.0,
                     0.0, 0.0, (const double*)nullptr, (const double*)nullptr,
                     (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                     (const double*)nullptr, (int*)nullptr, (const int*)nullptr, J);
  int P = 1;
  while (P < maxrk) P <<= 1;
  hipLaunchKernelGGL(k_unique<true>, dim3(1, nb), dim3(1024), sizeof(int) * (size_t)P, st,
                     (const int*)nullptr, (int*)nullptr, (int*)nullptr, J);
  hipLaunchKernelGGL(k_km_batch_clear_flags, dim3(std::min(gn, 1024u), nb), dim3(256), 0, st, J);
  hipLaunchKernelGGL(k_mark_centers<true>, dim3(cdiv(maxrk, 256), nb), dim3(256), 0, st,
                     (const int*)nullptr, (uint8_t*)nullptr, (int*)nullptr, (const int*)nullptr, J);
  hipLaunchKernelGGL(k_flag_count<true>, dim3(maxflag, nb), dim3(256), 0, st, (const uint8_t*)nullptr, 0,
                     0, (int*)nullptr, J);
  hipLaunchKernelGGL(k_scan_single<true>, dim3(1, nb), dim3(1024), 0, st, (int*)nullptr, 0,
                     (int*)nullptr, J);
  hipLaunchKernelGGL(k_flag_scatter<true>, dim3(maxflag, nb), dim3(256), 0, st, (const uint8_t*)nullptr,
                     0, 0, (const int*)nullptr, 0, (int*)nullptr, (const int*)nullptr, J);
  ISDF_HIP_LAUNCHED("k-means finish (batched)");
  return 0;
}

int kmeans_max_rk() { return kMaxRk; }

}  // namespace isdf_hip
