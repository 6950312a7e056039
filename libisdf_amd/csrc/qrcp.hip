// QRCP interpolation-point selection on the device: LibISDF's other point
// selection, getpoints_QRCP (src/cpu/getpoints_QRCP.cpp:11-500), at one rank.
//
// The reference sketches the pair products with two Gaussian matrices,
//   phiG = Psi_2 G2 (Ng x numPre),  psiG = Psi_1 G1 (Ng x numPre),
//   MG(i + j numPre, r) = phiG(r, i) psiG(r, j)      (getpoints_QRCP.cpp:200-232),
// keeps the grid columns whose squared norm exceeds 1e-5 of the largest
// (:280-290), and takes the pivots of a column-pivoted QR of that
// numPre^2 x ntotMG matrix (ScaLAPACK pdgeqpf, :440) as piv.
//
// Here MG is never formed.  Column j of MG is phiG(j) (x) psiG(j), so for the
// k-th orthonormal column q_k of the QR, R(k, j) = q_k^T MG(:, j) =
// phiG(j)^T Q_k psiG(j), with Q_k the numPre x numPre reshape of q_k.  Each of
// the first rk steps of the pivoted QR then is:
//   1. pivot: the largest partial column norm among positions >= k, the first
//      one in position order on ties, swapped into position k with its data
//      (pdgeqpf's idamax and column swap);
//   2. q_k: the pivot column of MG, orthogonalised against q_0..q_{k-1} by two
//      classical Gram-Schmidt passes and normalised (the Householder step's
//      new column of Q, up to sign);
//   3. R(k, j) for every position j > k from the bilinear form above, and the
//      partial norms downdated with pdgeqpf's formula; where that formula
//      asks for a fresh norm, it is taken as sqrt(|MG(:, j)|^2 - sum of R(., j)^2).
// piv[0..rk) are the grid indices of the first rk pivots; positions rk..
// keep the order the swaps left (the reference runs the factorization to
// numPre^2 steps, but only the first rk pivots enter ISDF).
//
// The Gaussian matrices follow the reference exactly: srand48(41)
// (src/common/isdf.cpp:107), drand48 pairs through Box-Muller
// (include/common/utility.hpp:455-471), G1 filled first, column-major.  The
// reference orthonormalises them (lapack::Orth, dgesvd); any orthonormal basis
// of the same column span gives the same pivots in exact arithmetic (MG's rows
// change by an orthogonal transformation), and a Householder QR is used here.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/isdf_hip.h"
#include "isdf_internal.h"
#include "mfma_gemm.h"

namespace isdf_hip {

namespace {

__host__ __device__ inline long long cdivq(long long a, long long b) { return (a + b - 1) / b; }

// POSIX drand48 with srand48(seed): X <- (0x5DEECE66D X + 0xB) mod 2^48,
// returned as X / 2^48.  Kept local so the caller's own drand48 state is left
// alone.
struct Drand48 {
  uint64_t x;
  explicit Drand48(long seed) : x(((uint64_t)(uint32_t)seed << 16) | 0x330EULL) {}
  double next() {
    x = (0x5DEECE66DULL * x + 0xBULL) & ((1ULL << 48) - 1);
    return (double)x / 281474976710656.0;
  }
};

// Orthonormal basis of the columns of G (m x n, column-major, m >= n), in
// place: Householder QR, then Q formed explicitly.
void orth(int m, int n, std::vector<double>& G) {
  std::vector<double> V((size_t)m * n, 0.0), tau((size_t)n, 0.0);
  std::vector<double> A = G;
  for (int k = 0; k < n; k++) {
    double* a = &A[(size_t)k * m];
    double s = 0.0;
    for (int i = k; i < m; i++) s += a[i] * a[i];
    const double nrm = std::sqrt(s);
    double* v = &V[(size_t)k * m];
    if (nrm == 0.0) continue;
    const double alpha = a[k] >= 0 ? -nrm : nrm;
    for (int i = k; i < m; i++) v[i] = a[i];
    v[k] -= alpha;
    double vv = 0.0;
    for (int i = k; i < m; i++) vv += v[i] * v[i];
    tau[k] = vv > 0 ? 2.0 / vv : 0.0;
    for (int j = k; j < n; j++) {
      double* c = &A[(size_t)j * m];
      double d = 0.0;
      for (int i = k; i < m; i++) d += v[i] * c[i];
      d *= tau[k];
      for (int i = k; i < m; i++) c[i] -= d * v[i];
    }
  }
  // Q = H_0 H_1 ... H_{n-1} applied to the first n columns of I.
  for (int j = 0; j < n; j++) {
    double* q = &G[(size_t)j * m];
    for (int i = 0; i < m; i++) q[i] = (i == j) ? 1.0 : 0.0;
    for (int k = n - 1; k >= 0; k--) {
      const double* v = &V[(size_t)k * m];
      double d = 0.0;
      for (int i = k; i < m; i++) d += v[i] * q[i];
      d *= tau[k];
      for (int i = k; i < m; i++) q[i] -= d * v[i];
    }
  }
}

// out[i * ng + r] = sum_k psi[(s + k) ld + r] G[k + i K]  (psi rows times G).
__global__ __launch_bounds__(256) void k_qr_proj(const double* __restrict__ psi, long long ld,
                                                 int ng, int s, int K,
                                                 const double* __restrict__ G, int np,
                                                 double* __restrict__ out) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)np * ng) return;
  const int i = (int)(e / ng), r = (int)(e % ng);
  double acc = 0.0;
  for (int k = 0; k < K; k++) acc += psi[(long long)(s + k) * ld + r] * G[k + (long long)i * K];
  out[e] = acc;
}

// Squared MG column norms |phiG(r)|^2 |psiG(r)|^2 and their maximum (blocks
// write partial maxima).
__global__ __launch_bounds__(256) void k_qr_norms(const double* __restrict__ PH,
                                                  const double* __restrict__ PS, int ng, int np,
                                                  double* __restrict__ n2,
                                                  double* __restrict__ pmax) {
  __shared__ double sm[256];
  const int r = blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (r < ng) {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < np; i++) {
      const double x = PH[(long long)i * ng + r], y = PS[(long long)i * ng + r];
      a += x * x;
      b += y * y;
    }
    v = a * b;
    n2[r] = v;
  }
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) pmax[blockIdx.x] = sm[0];
}

// One workgroup: the grid columns above 1e-5 of the largest squared norm
// (getpoints_QRCP.cpp:280-290), in ascending grid order: perm[pos] = grid
// index; scal[0] = their count.
__global__ __launch_bounds__(1024) void k_qr_compact(const double* __restrict__ n2,
                                                     const double* __restrict__ pmax, int nblk,
                                                     int ng, int* __restrict__ perm,
                                                     int* __restrict__ scal) {
  __shared__ double smax;
  __shared__ int wsum[16], base;
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int b = 0; b < nblk; b++) m = fmax(m, pmax[b]);
    smax = m;
    base = 0;
  }
  __syncthreads();
  const double tol = smax * 1e-5;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int r0 = 0; r0 < ng; r0 += 1024) {
    const int r = r0 + threadIdx.x;
    const bool keep = r < ng && n2[r] > tol;
    const unsigned long long bal = __ballot(keep);
    const int before = __popcll(bal & ((1ULL << lane) - 1ULL));
    if (lane == 0) wsum[wave] = __popcll(bal);
    __syncthreads();
    int off = base;
    for (int w = 0; w < wave; w++) off += wsum[w];
    if (keep) perm[off + before] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int w = 0; w < 16; w++) t += wsum[w];
      base += t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) scal[0] = base;
}

// The kept columns in position order: sketch rows, vn1 = vn2 = the column
// norm, a0n2 its square, s2 (sum of R(., j)^2 so far) = 0.
__global__ __launch_bounds__(256) void k_qr_gather(const int* __restrict__ perm,
                                                   const int* __restrict__ scal, int ng, int np,
                                                   const double* __restrict__ PH,
                                                   const double* __restrict__ PS,
                                                   const double* __restrict__ n2,
                                                   double* __restrict__ PHc,
                                                   double* __restrict__ PSc,
                                                   double* __restrict__ vn1,
                                                   double* __restrict__ vn2,
                                                   double* __restrict__ a0n2,
                                                   double* __restrict__ s2) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  const int n = scal[0];
  if (e >= (long long)np * ng) return;
  const int i = (int)(e / ng), pos = (int)(e % ng);
  if (pos >= n) return;
  const int r = perm[pos];
  PHc[e] = PH[(long long)i * ng + r];
  PSc[e] = PS[(long long)i * ng + r];
  if (i == 0) {
    const double v = sqrt(n2[r]);
    vn1[pos] = v;
    vn2[pos] = v;
    a0n2[pos] = n2[r];
    s2[pos] = 0.0;
  }
}

// Partial argmax of vn1 over positions [lo, n): per block (value, position),
// first position on ties.
__device__ inline void argmax_pair(double& v, int& p, double v2, int p2) {
  if (v2 > v || (v2 == v && p2 < p)) {
    v = v2;
    p = p2;
  }
}

__global__ __launch_bounds__(256) void k_qr_argmax(const double* __restrict__ vn1, int lo, int n,
                                                   double* __restrict__ pv, int* __restrict__ pp) {
  __shared__ double sv[256];
  __shared__ int sp[256];
  const int j = lo + blockIdx.x * 256 + threadIdx.x;
  double v = -1.0;
  int p = 0x7fffffff;
  if (j < n) {
    v = vn1[j];
    p = j;
  }
  sv[threadIdx.x] = v;
  sp[threadIdx.x] = p;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) {
      double a = sv[threadIdx.x];
      int b = sp[threadIdx.x];
      argmax_pair(a, b, sv[threadIdx.x + h], sp[threadIdx.x + h]);
      sv[threadIdx.x] = a;
      sp[threadIdx.x] = b;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pp[blockIdx.x] = sp[0];
  }
}

// Step k, part 1 (one workgroup): the pivot from the partial maxima, swapped
// into position k with its data, and the pivot column of MG into a.
__global__ __launch_bounds__(256) void k_qr_pivot(int k, int nparts, const double* __restrict__ pv,
                                                  const int* __restrict__ pp, int ldp, int np,
                                                  int* __restrict__ perm, double* __restrict__ PHc,
                                                  double* __restrict__ PSc,
                                                  double* __restrict__ vn1,
                                                  double* __restrict__ vn2,
                                                  double* __restrict__ a0n2,
                                                  double* __restrict__ s2,
                                                  double* __restrict__ a) {
  __shared__ double sv[256];
  __shared__ int sp[256];
  {
    double v = -1.0;
    int q = 0x7fffffff;
    for (int b = threadIdx.x; b < nparts; b += 256) argmax_pair(v, q, pv[b], pp[b]);
    sv[threadIdx.x] = v;
    sp[threadIdx.x] = q;
  }
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) {
      double v = sv[threadIdx.x];
      int q = sp[threadIdx.x];
      argmax_pair(v, q, sv[threadIdx.x + h], sp[threadIdx.x + h]);
      sv[threadIdx.x] = v;
      sp[threadIdx.x] = q;
    }
    __syncthreads();
  }
  const int p = sp[0];
  if (p != k) {
    if (threadIdx.x == 0) {
      int t = perm[p];
      perm[p] = perm[k];
      perm[k] = t;
      // pdgeqpf: the pivot's norms move to k; position p takes k's.
      double x = vn1[p];
      vn1[p] = vn1[k];
      vn1[k] = x;
      x = vn2[p];
      vn2[p] = vn2[k];
      vn2[k] = x;
      x = a0n2[p];
      a0n2[p] = a0n2[k];
      a0n2[k] = x;
      x = s2[p];
      s2[p] = s2[k];
      s2[k] = x;
    }
    for (int i = threadIdx.x; i < np; i += 256) {
      double x = PHc[(long long)i * ldp + p];
      PHc[(long long)i * ldp + p] = PHc[(long long)i * ldp + k];
      PHc[(long long)i * ldp + k] = x;
      x = PSc[(long long)i * ldp + p];
      PSc[(long long)i * ldp + p] = PSc[(long long)i * ldp + k];
      PSc[(long long)i * ldp + k] = x;
    }
  }
  __syncthreads();
  // MG(i + j np, k) = phiG(k, i) psiG(k, j).
  const int m = np * np;
  for (int e = threadIdx.x; e < m; e += 256) {
    const int i = e % np, j = e / np;
    a[e] = PHc[(long long)i * ldp + k] * PSc[(long long)j * ldp + k];
  }
}

// c[t] = q_t^T a, t < k: one wave per t.
__global__ __launch_bounds__(256) void k_qr_dots(const double* __restrict__ Q, int m, int k,
                                                 const double* __restrict__ a,
                                                 double* __restrict__ c) {
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (t >= k) return;
  const double* q = Q + (long long)t * m;
  double s = 0.0;
  for (int i = lane; i < m; i += 64) s += q[i] * a[i];
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) c[t] = s;
}

// a[i] -= sum_t Q[i, t] c[t], t < k: one wave per row i of the row-major copy
// Qr[i * ldr + t] (lanes over t, then a fixed butterfly).
__global__ __launch_bounds__(256) void k_qr_axpy(const double* __restrict__ Qr, int ldr, int m,
                                                 int k, const double* __restrict__ c,
                                                 double* __restrict__ a) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= m) return;
  const double* q = Qr + (long long)i * ldr;
  double s = 0.0;
  for (int t = lane; t < k; t += 64) s += q[t] * c[t];
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) a[i] -= s;
}

// Step k, part 3: q_k = a / |a| (every workgroup forms it in LDS; workgroup 0
// stores it as column k of Q), then for positions j in (k, n) of this
// workgroup's tile of T: R(k, j) = sum_jj psiG(j, jj) sum_i q(i + jj np) phiG(j, i),
// pdgeqpf's norm downdate, and the tile's partial argmax for step k + 1.
__global__ void k_qr_update(int k, int n, int np, int ldp, const double* __restrict__ a,
                            double* __restrict__ Q, double* __restrict__ Qr, int ldr,
                            const double* __restrict__ PHc,
                            const double* __restrict__ PSc, double* __restrict__ vn1,
                            double* __restrict__ vn2, const double* __restrict__ a0n2,
                            double* __restrict__ s2, double* __restrict__ pv,
                            int* __restrict__ pp) {
  extern __shared__ double lds[];
  const int T = blockDim.x, m = np * np;
  double* q = lds;          // m
  double* ph = lds + m;     // np x T
  __shared__ double red[256];
  __shared__ int redp[256];
  // |a|^2 in a fixed order: per thread strided, then a fixed tree.
  double s = 0.0;
  for (int e = threadIdx.x; e < m; e += T) s += a[e] * a[e];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = T / 2; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double nrm = sqrt(red[0]);
  const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
  __syncthreads();
  for (int e = threadIdx.x; e < m; e += T) {
    const double v = a[e] * inv;
    q[e] = v;
    if (blockIdx.x == 0) {
      Q[(long long)k * m + e] = v;
      Qr[(long long)e * ldr + k] = v;
    }
  }
  const int j0 = k + 1 + blockIdx.x * T;
  const int j = j0 + threadIdx.x;
  for (int i = 0; i < np; i++)
    ph[i * T + threadIdx.x] = (j < n) ? PHc[(long long)i * ldp + j] : 0.0;
  __syncthreads();
  double v = -1.0;
  int p = 0x7fffffff;
  if (j < n) {
    double y = 0.0;
    for (int jj = 0; jj < np; jj++) {
      double u = 0.0;
      const double* qc = q + jj * np;
      for (int i = 0; i < np; i++) u += qc[i] * ph[i * T + threadIdx.x];
      y += u * PSc[(long long)jj * ldp + j];
    }
    const double s2n = s2[j] + y * y;
    s2[j] = s2n;
    double w1 = vn1[j];
    if (w1 != 0.0) {
      double temp = fabs(y) / w1;
      temp = 1.0 - temp * temp;
      temp = temp > 0.0 ? temp : 0.0;
      const double ratio = w1 / vn2[j];
      const double temp2 = 1.0 + 0.05 * temp * ratio * ratio;
      if (temp2 == 1.0) {
        const double rest = a0n2[j] - s2n;
        w1 = rest > 0.0 ? sqrt(rest) : 0.0;
        vn2[j] = w1;
      } else {
        w1 = w1 * sqrt(temp);
      }
      vn1[j] = w1;
    }
    v = w1;
    p = j;
  }
  red[threadIdx.x] = v;
  redp[threadIdx.x] = p;
  __syncthreads();
  for (int h = T / 2; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) {
      double a1 = red[threadIdx.x];
      int b1 = redp[threadIdx.x];
      argmax_pair(a1, b1, red[threadIdx.x + h], redp[threadIdx.x + h]);
      red[threadIdx.x] = a1;
      redp[threadIdx.x] = b1;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = red[0];
    pp[blockIdx.x] = redp[0];
  }
}

// The same step on FP64 MFMA: for blocks of 64 positions (16 per wave; each
// workgroup walks kQrBlocks blocks, so q is set up once per workgroup),
// T(j, i) = sum_jj psiG(j, jj) q(i + jj np) as 16 x 16 x 4 products (A = the
// psiG rows, B = q reshaped), then R(k, j) = sum_i phiG(j, i) T(j, i) with the
// phiG tile staged in LDS and the 16 columns of a tile summed across their
// lanes.  Used when q and the phiG tile fit 60 KB of LDS.
constexpr int kQrBlocks = 4;
__global__ __launch_bounds__(256) void k_qr_update_mfma(
    int k, int n, int np, int ldp, const double* __restrict__ a, double* __restrict__ Q,
    double* __restrict__ Qr, int ldr, const double* __restrict__ PHc,
    const double* __restrict__ PSc, double* __restrict__ vn1, double* __restrict__ vn2,
    const double* __restrict__ a0n2, double* __restrict__ s2, double* __restrict__ pv,
    int* __restrict__ pp) {
  extern __shared__ double lds[];
  const int m = np * np;
  double* q = lds;       // m
  double* ph = lds + m;  // np x 64: ph[i * 64 + t]
  __shared__ double red[256];
  __shared__ int redp[256];
  double s = 0.0;
  for (int e = threadIdx.x; e < m; e += 256) s += a[e] * a[e];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double nrm = sqrt(red[0]);
  const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
  __syncthreads();
  for (int e = threadIdx.x; e < m; e += 256) {
    const double v = a[e] * inv;
    q[e] = v;
    if (blockIdx.x == 0) {
      Q[(long long)k * m + e] = v;
      Qr[(long long)e * ldr + k] = v;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  double best = -1.0;
  int bestp = 0x7fffffff;
  for (int blk = 0; blk < kQrBlocks; blk++) {
    const int j0 = k + 1 + (blockIdx.x * kQrBlocks + blk) * 64;
    if (j0 >= n) break;  // uniform across the workgroup
    __syncthreads();     // q written (first block) / previous tile consumed
    for (int e = threadIdx.x; e < np * 64; e += 256) {
      const int i = e / 64, j = j0 + e % 64;
      ph[e] = (j < n) ? PHc[(long long)i * ldp + j] : 0.0;
    }
    __syncthreads();
    const int p0 = j0 + 16 * wave;
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    for (int n0 = 0; n0 < np; n0 += 16) {
      d4 acc = {0.0, 0.0, 0.0, 0.0};
      const int ib = n0 + lr;
      for (int k0 = 0; k0 < np; k0 += 4) {
        const int jj = k0 + lg;
        const int pa = p0 + lr;
        const double av = (jj < np && pa < n) ? PSc[(long long)jj * ldp + pa] : 0.0;
        const double bv = (jj < np && ib < np) ? q[ib + jj * np] : 0.0;
        acc = mfma_f64(av, bv, acc);
      }
      if (ib < np) {
#pragma unroll
        for (int r = 0; r < 4; r++) part[r] += acc[r] * ph[ib * 64 + 16 * wave + lg + 4 * r];
      }
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
      for (int off = 8; off >= 1; off >>= 1) part[r] += __shfl_xor(part[r], off, 64);
    const int j = p0 + lg + 4 * lr;
    if (lr < 4 && j < n) {
      const double y = lr == 0 ? part[0] : lr == 1 ? part[1] : lr == 2 ? part[2] : part[3];
      const double s2n = s2[j] + y * y;
      s2[j] = s2n;
      double w1 = vn1[j];
      if (w1 != 0.0) {
        double temp = fabs(y) / w1;
        temp = 1.0 - temp * temp;
        temp = temp > 0.0 ? temp : 0.0;
        const double ratio = w1 / vn2[j];
        const double temp2 = 1.0 + 0.05 * temp * ratio * ratio;
        if (temp2 == 1.0) {
          const double rest = a0n2[j] - s2n;
          w1 = rest > 0.0 ? sqrt(rest) : 0.0;
          vn2[j] = w1;
        } else {
          w1 = w1 * sqrt(temp);
        }
        vn1[j] = w1;
      }
      argmax_pair(best, bestp, w1, j);
    }
  }
  red[threadIdx.x] = best;
  redp[threadIdx.x] = bestp;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) {
      double a1 = red[threadIdx.x];
      int b1 = redp[threadIdx.x];
      argmax_pair(a1, b1, red[threadIdx.x + h], redp[threadIdx.x + h]);
      red[threadIdx.x] = a1;
      redp[threadIdx.x] = b1;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = red[0];
    pp[blockIdx.x] = redp[0];
  }
}

// piv[j] = perm[j] for j < n, 0 beyond (the reference zeroes pivQR_ first).
__global__ __launch_bounds__(256) void k_qr_piv(const int* __restrict__ perm,
                                                const int* __restrict__ scal, int ng,
                                                int* __restrict__ piv) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j < ng) piv[j] = j < scal[0] ? perm[j] : 0;
}

int iround(double a) {
  if (a > 0) return (a - (int)a < 0.5) ? (int)a : ((int)a + 1);
  return ((int)a - a < 0.5) ? (int)a : ((int)a - 1);
}

}  // namespace

int qrcp_num_pre(int rk) { return iround(std::sqrt(rk * 1.5)); }

size_t qrcp_bytes(long long ng, int rk) {
  const long long np = qrcp_num_pre(rk), m = np * np;
  const long long nblk = cdivq(ng, 256);
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const long long nparts = cdivq(ng, 32);  // partial maxima: tiles of at least 32 positions
  return al(8 * (size_t)(np * ng)) * 4 + al(8 * (size_t)ng) * 5 + al(4 * (size_t)ng) +
         al(8 * (size_t)(m * rk)) * 2 + al(8 * (size_t)m) + al(8 * (size_t)rk) +
         al(8 * (size_t)nblk) + al(8 * (size_t)nparts) + al(4 * (size_t)nparts) + 256;
}

// The QRCP selection into piv_dev (ng ints).  Returns an ISDF status; on
// argument errors nothing is launched.
int qrcp_select(const isdf_problem* p, const double* d_psi, long long ld, int ng, int rk,
                void* scratch, int* piv_dev, hipStream_t st) {
  const int np = qrcp_num_pre(rk), m = np * np;
  const int s1 = p->nocc - p->nv1, K1 = p->nv1 + p->nc1;
  const int s2 = p->nocc - p->nv2, K2 = p->nv2 + p->nc2;
  // getpoints_QRCP.cpp:80-84 and :425-428.
  if (np > p->nstate)
    return set_error(ISDF_EINVAL, "numMu is too large for interpolative separable density fitting!");
  if ((long long)m > ng)
    return set_error(ISDF_EINVAL,
                     "numPre*numPre > ntot. The number of grid points is perhaps too small!");
  if (np > K1 || np > K2)
    return set_error(ISDF_EINVAL, "QRCP: numPre = " + std::to_string(np) +
                                      " exceeds a pair set's size (nv + nc)");
  if (m < rk) return set_error(ISDF_EINVAL, "QRCP: numPre^2 < rk");

  // Gaussian sketches, G1 then G2, then orthonormalised.
  Drand48 rng(41);
  std::vector<double> G1((size_t)K1 * np), G2((size_t)K2 * np);
  const double kPi = 3.14159265358979323846;
  for (auto* G : {&G1, &G2})
    for (double& x : *G) {
      const double u = rng.next(), v = rng.next();
      x = std::sqrt(-2.0 * std::log(u)) * std::cos(2.0 * kPi * v);
    }
  orth(K1, np, G1);
  orth(K2, np, G2);

  const long long nblk = cdivq(ng, 256);
  char* b = (char*)scratch;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* r = b + off;
    off += (bytes + 255) & ~(size_t)255;
    return r;
  };
  double* PH = (double*)take(8 * (size_t)np * ng);
  double* PS = (double*)take(8 * (size_t)np * ng);
  double* PHc = (double*)take(8 * (size_t)np * ng);
  double* PSc = (double*)take(8 * (size_t)np * ng);
  double* n2 = (double*)take(8 * (size_t)ng);
  double* vn1 = (double*)take(8 * (size_t)ng);
  double* vn2 = (double*)take(8 * (size_t)ng);
  double* a0n2 = (double*)take(8 * (size_t)ng);
  double* s2v = (double*)take(8 * (size_t)ng);
  int* perm = (int*)take(4 * (size_t)ng);
  double* Q = (double*)take(8 * (size_t)m * rk);
  double* Qr = (double*)take(8 * (size_t)m * rk);  // the same columns, row-major (ld rk)
  double* av = (double*)take(8 * (size_t)m);
  double* cv = (double*)take(8 * (size_t)rk);
  // Partial maxima: one per block of the argmax (256 positions) or of the
  // update (T >= 32 positions).
  const long long nparts_max = cdivq(ng, 32);
  double* pmax = (double*)take(8 * (size_t)nblk);
  double* pv = (double*)take(8 * (size_t)nparts_max);
  int* pp = (int*)take(4 * (size_t)nparts_max);
  int* scal = (int*)take(256);
  // G1 | G2 are staged in Q's buffer, which the steps fill only later
  // (np <= K1, K2 <= nstate and m >= rk, so (K1 + K2) np <= 2 nstate np; checked).
  double* g1 = Q;
  double* g2 = Q + (size_t)K1 * np;
  if ((size_t)(K1 + K2) * np > (size_t)m * rk)
    return set_error(ISDF_EINVAL, "QRCP: sketch does not fit the Q buffer");
  ISDF_HIP_TRY(hipMemcpyAsync(g1, G1.data(), 8 * G1.size(), hipMemcpyHostToDevice, st));
  ISDF_HIP_TRY(hipMemcpyAsync(g2, G2.data(), 8 * G2.size(), hipMemcpyHostToDevice, st));
  const dim3 gp((unsigned)cdivq((long long)np * ng, 256));
  // phiG = Psi_2 G2, psiG = Psi_1 G1 (getpoints_QRCP.cpp:200-208).
  hipLaunchKernelGGL(k_qr_proj, gp, dim3(256), 0, st, d_psi, ld, ng, s2, K2, (const double*)g2,
                     np, PH);
  hipLaunchKernelGGL(k_qr_proj, gp, dim3(256), 0, st, d_psi, ld, ng, s1, K1, (const double*)g1,
                     np, PS);
  hipLaunchKernelGGL(k_qr_norms, dim3((unsigned)nblk), dim3(256), 0, st, (const double*)PH,
                     (const double*)PS, ng, np, n2, pmax);
  hipLaunchKernelGGL(k_qr_compact, dim3(1), dim3(1024), 0, st, (const double*)n2,
                     (const double*)pmax, (int)nblk, ng, perm, scal);
  hipLaunchKernelGGL(k_qr_gather, gp, dim3(256), 0, st, (const int*)perm, (const int*)scal, ng,
                     np, (const double*)PH, (const double*)PS, (const double*)n2, PHc, PSc, vn1,
                     vn2, a0n2, s2v);
  ISDF_HIP_LAUNCHED("QRCP setup");
  int nmg = 0;
  ISDF_HIP_TRY(hipMemcpyAsync(&nmg, scal, sizeof(int), hipMemcpyDeviceToHost, st));
  ISDF_HIP_TRY(hipStreamSynchronize(st));
  if (nmg < rk)
    return set_error(ISDF_EINVAL, "QRCP: only " + std::to_string(nmg) +
                                      " grid columns above the norm threshold, fewer than rk");
  // Tile of the update kernel: positions per workgroup, with q and the phiG
  // tile in LDS.
  // Dynamic LDS stays within the 64 KB a launch gets without attributes.
  int T = 128;
  while (T > 32 && 8 * ((size_t)m + (size_t)np * T) > 60 * 1024) T /= 2;
  const size_t lds = 8 * ((size_t)m + (size_t)np * T);
  if (lds > 60 * 1024)
    return set_error(ISDF_EINVAL, "QRCP: numPre = " + std::to_string(np) +
                                      " is above the supported 80 (rk above ~4300)");
  const size_t lds_mfma = 8 * ((size_t)m + (size_t)np * 64);
  const bool use_mfma = lds_mfma <= 60 * 1024 && std::getenv("ISDF_QRCP_VALU") == nullptr;
  int nparts = (int)cdivq(nmg, 256);
  hipLaunchKernelGGL(k_qr_argmax, dim3((unsigned)nparts), dim3(256), 0, st, (const double*)vn1,
                     0, nmg, pv, pp);
  for (int k = 0; k < rk; k++) {
    hipLaunchKernelGGL(k_qr_pivot, dim3(1), dim3(256), 0, st, k, nparts, (const double*)pv,
                       (const int*)pp, ng, np, perm, PHc, PSc, vn1, vn2, a0n2, s2v, av);
    if (k > 0) {
      for (int pass = 0; pass < 2; pass++) {
        hipLaunchKernelGGL(k_qr_dots, dim3((unsigned)cdivq(k, 4)), dim3(256), 0, st,
                           (const double*)Q, m, k, (const double*)av, cv);
        hipLaunchKernelGGL(k_qr_axpy, dim3((unsigned)cdivq(m, 4)), dim3(256), 0, st,
                           (const double*)Qr, rk, m, k, (const double*)cv, av);
      }
    }
    const int rest = nmg - (k + 1);
    if (use_mfma) {
      const int nb = std::max(1, (int)cdivq(rest, 64 * kQrBlocks));
      hipLaunchKernelGGL(k_qr_update_mfma, dim3((unsigned)nb), dim3(256), lds_mfma, st, k, nmg,
                         np, ng, (const double*)av, Q, Qr, rk, (const double*)PHc,
                         (const double*)PSc, vn1, vn2, (const double*)a0n2, s2v, pv, pp);
      nparts = nb;
    } else {
      const int nb = std::max(1, (int)cdivq(rest, T));
      hipLaunchKernelGGL(k_qr_update, dim3((unsigned)nb), dim3(T), lds, st, k, nmg, np, ng,
                         (const double*)av, Q, Qr, rk, (const double*)PHc, (const double*)PSc,
                         vn1, vn2, (const double*)a0n2, s2v, pv, pp);
      nparts = nb;
    }
    ISDF_HIP_LAUNCHED("QRCP step");
  }
  hipLaunchKernelGGL(k_qr_piv, dim3((unsigned)nblk), dim3(256), 0, st, (const int*)perm,
                     (const int*)scal, ng, piv_dev);
  ISDF_HIP_LAUNCHED("k_qr_piv");
  return ISDF_OK;
}

}  // namespace isdf_hip
