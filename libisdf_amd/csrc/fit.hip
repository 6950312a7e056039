// Fit error of an ISDF result on the device: the fit-error half of LibISDF's
// check=true validator (src/common/checkcode.cpp:62-166),
//
//   fit = ||Z - Theta C||_F / ||Z||_F,   Z(r, (k,j)) = psi_k(r) psi_j(r),
//                                        C(mu, (k,j)) = psi_k(r_mu) psi_j(r_mu),
//
// over all Ng grid rows and all K1*K2 pairs (k in S1, j in S2; pair index
// q = k*K2 + j as checkcode orders it).  Z is never stored: the residual is
// formed in the epilogue of an FP64 MFMA GEMM Theta (Ng x rk) * C (rk x K1K2),
// where each accumulator element meets its Z entry, recomputed from two psi
// values, and only the two sums of squares leave the tile.  C is built per slab
// of pair columns (at most kSlabBytes), so memory stays O(rk * slab) whatever
// K1*K2 is.  Every sum has a fixed order (per tile, then per slab into the same
// partial slot, then one workgroup over the partials), so the result is
// deterministic.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/isdf_hip.h"
#include "isdf_internal.h"
#include "mfma_gemm.h"

namespace isdf_hip {

int problem_rank(const isdf_problem* p);
int validate_problem(const isdf_problem* p, long long* ng_out);
std::mutex& solve_mutex();

namespace {

__host__ __device__ inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

constexpr size_t kSlabBytes = 32u << 20;  // C slab: rk x nq doubles

// C(mu, q) for the slab q in [q0, q0+nq): Cs[(q-q0)*rk + mu] (mu contiguous).
__global__ __launch_bounds__(256) void k_fit_cmat(const double* __restrict__ psi, long long ld,
                                                  const int* __restrict__ piv, int rk, int s1,
                                                  int s2, int K2, long long q0, int nq,
                                                  double* __restrict__ Cs) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)nq * rk) return;
  const int mu = (int)(e % rk), qq = (int)(e / rk);
  const long long q = q0 + qq;
  const int k = (int)(q / K2), j = (int)(q % K2);
  const long long g = piv[mu];
  Cs[e] = psi[(long long)(s1 + k) * ld + g] * psi[(long long)(s2 + j) * ld + g];
}

// One (row strip, pair tile) of the residual GEMM.  part[2b], part[2b+1]
// receive (first slab) or accumulate (later slabs) the tile's sum of squared
// residuals and of squared Z entries.
typedef MfmaTile<256, 128, 16, 4, 2, false> FitTile;  // 8 waves of 64 x 64

__device__ __forceinline__ void xcd_strip(int b, int ntile, int nstrip, int* s, int* t) {
  // The ntile pair tiles of one row strip land on one XCD (workgroups are
  // dealt round-robin to the 8 XCDs) and run back to back, so the strip of
  // Theta is read from HBM once into that XCD's L2.
  const int full = (nstrip / 8) * 8;
  if (b < full * ntile) {
    const int x = b & 7, l = b >> 3;
    *s = x + 8 * (l / ntile);
    *t = l % ntile;
  } else {
    const int b2 = b - full * ntile;
    *s = full + b2 / ntile;
    *t = b2 % ntile;
  }
}

__global__ __launch_bounds__(FitTile::NT, 1) void k_fit_resid(
    int ng, int rk, const double* __restrict__ theta, const double* __restrict__ Cs, int nq,
    const double* __restrict__ psi, long long ld, int s1, int s2, int K2, long long q0,
    int first, double* __restrict__ part) {
  __shared__ double lds[FitTile::LDS_DOUBLES];
  __shared__ double red[2][FitTile::NT / 64];
  const int nstrip = cdiv(ng, FitTile::BM);
  const int ntile = cdiv(nq, FitTile::BN);
  int s, t;
  xcd_strip(blockIdx.x, ntile, nstrip, &s, &t);
  const long long m0 = (long long)s * FitTile::BM;
  const int n0 = t * FitTile::BN;
  const int mlim = (int)std::min<long long>(FitTile::BM, ng - m0);
  const int nlim = std::min(FitTile::BN, nq - n0);
  typename FitTile::Acc acc;
  FitTile::zero(acc);
  // A(m, k) = Theta(m0+m, k) at theta[k*ng + m0 + m]; B(k, n) = Cs[(n0+n)*rk + k].
  FitTile::run(theta + m0, ng, Cs + (long long)n0 * rk, rk, rk, mlim, nlim, lds, acc);
  double e = 0.0, z = 0.0;
#pragma unroll
  for (int jj = 0; jj < FitTile::TN; jj++) {
    const int n = FitTile::col(jj);
    if (n < nlim) {
      const long long q = q0 + n0 + n;
      const int k = (int)(q / K2), j = (int)(q % K2);
      const double* pk = psi + (long long)(s1 + k) * ld + m0;
      const double* pj = psi + (long long)(s2 + j) * ld + m0;
#pragma unroll
      for (int i = 0; i < FitTile::TM; i++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int m = FitTile::row(i, r);
          if (m < mlim) {
            const double zz = pk[m] * pj[m];
            const double d = zz - acc[i][jj][r];
            e += d * d;
            z += zz * zz;
          }
        }
    }
  }
  // Wave sums (butterfly over the 64 lanes), then the 8 waves in order.
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    e += __shfl_xor(e, off, 64);
    z += __shfl_xor(z, off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = e;
    red[1][wave] = z;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double se = 0.0, sz = 0.0;
    for (int w = 0; w < FitTile::NT / 64; w++) {
      se += red[0][w];
      sz += red[1][w];
    }
    double* p = part + 2 * (long long)blockIdx.x;
    if (first) {
      p[0] = se;
      p[1] = sz;
    } else {
      p[0] += se;
      p[1] += sz;
    }
  }
}

// One workgroup: out[0] = sum of part[2b], out[1] = sum of part[2b+1], each
// thread over a fixed strided subset, then a fixed tree.
__global__ __launch_bounds__(1024) void k_fit_total(const double* __restrict__ part, int nb,
                                                    double* __restrict__ out) {
  __shared__ double se[1024], sz[1024];
  double e = 0.0, z = 0.0;
  for (int b = threadIdx.x; b < nb; b += 1024) {
    e += part[2 * (long long)b];
    z += part[2 * (long long)b + 1];
  }
  se[threadIdx.x] = e;
  sz[threadIdx.x] = z;
  __syncthreads();
  for (int h = 512; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) {
      se[threadIdx.x] += se[threadIdx.x + h];
      sz[threadIdx.x] += sz[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = se[0];
    out[1] = sz[0];
  }
}

// Grow-only device buffer of the fit error, one per device.
struct FitBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};
std::vector<FitBuf> g_fit;

int fit_buffer(size_t bytes, void** out) {
  int dev = 0;
  ISDF_HIP_TRY(hipGetDevice(&dev));
  if ((int)g_fit.size() <= dev) g_fit.resize(dev + 1);
  FitBuf& b = g_fit[dev];
  if (b.bytes < bytes) {
    if (b.ptr) ISDF_HIP_TRY(hipFree(b.ptr));
    b.ptr = nullptr;
    b.bytes = 0;
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      b.ptr = nullptr;
      return set_error(ISDF_ENOMEM, "fit-error buffer allocation of " + std::to_string(bytes) +
                                        " bytes failed: " + hipGetErrorString(e));
    }
    b.bytes = bytes;
  }
  *out = b.ptr;
  return 0;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

// The fit error; the caller holds the library lock.
int fit_error_unlocked(const isdf_problem* p, const double* d_psi, const double* d_theta,
                       const int* h_piv, hipStream_t st, isdf_check* out) {
  long long ng = 0;
  if (int rc = validate_problem(p, &ng)) return rc;
  if (!d_psi || !d_theta || !h_piv || !out) return set_error(ISDF_EINVAL, "NULL argument");
  const int rk = problem_rank(p);
  for (int mu = 0; mu < rk; mu++)
    if (h_piv[mu] < 0 || h_piv[mu] >= ng)
      return set_error(ISDF_EINVAL, "piv[" + std::to_string(mu) + "] is not a grid index");
  const long long ld = p->ld_psi ? p->ld_psi : ng;
  const int s1 = p->nocc - p->nv1, K1 = p->nv1 + p->nc1;
  const int s2 = p->nocc - p->nv2, K2 = p->nv2 + p->nc2;
  const long long npair = (long long)K1 * K2;
  // Slab width: a multiple of the pair tile, at least one tile.
  long long nq = (long long)(kSlabBytes / (sizeof(double) * (size_t)rk));
  if (const char* v = std::getenv("ISDF_FIT_SLAB_COLS")) nq = std::atoll(v);
  nq = std::max<long long>(FitTile::BN, nq / FitTile::BN * FitTile::BN);
  nq = std::min<long long>(nq, (npair + FitTile::BN - 1) / FitTile::BN * FitTile::BN);
  const int nstrip = cdiv(ng, FitTile::BM);
  const int ntile_max = cdiv(nq, FitTile::BN);
  const long long nb = (long long)nstrip * ntile_max;
  if (nb > 0x7fffffffLL) return set_error(ISDF_EINVAL, "fit error: too many tiles");

  const size_t c_bytes = align256(sizeof(double) * (size_t)rk * nq);
  const size_t piv_bytes = align256(sizeof(int) * (size_t)rk);
  const size_t part_bytes = align256(2 * sizeof(double) * (size_t)nb);
  void* base = nullptr;
  if (int rc = fit_buffer(c_bytes + piv_bytes + part_bytes + 256, &base)) return rc;
  double* Cs = (double*)base;
  int* d_piv = (int*)((char*)base + c_bytes);
  double* part = (double*)((char*)base + c_bytes + piv_bytes);
  double* d_out = (double*)((char*)base + c_bytes + piv_bytes + part_bytes);

  ISDF_HIP_TRY(hipMemcpyAsync(d_piv, h_piv, sizeof(int) * rk, hipMemcpyHostToDevice, st));
  int first = 1;
  int nb_used = 0;
  for (long long q0 = 0; q0 < npair; q0 += nq) {
    const int w = (int)std::min<long long>(nq, npair - q0);
    hipLaunchKernelGGL(k_fit_cmat, dim3(cdiv((long long)w * rk, 256)), dim3(256), 0, st, d_psi,
                       ld, (const int*)d_piv, rk, s1, s2, K2, q0, w, Cs);
    ISDF_HIP_LAUNCHED("k_fit_cmat");
    // Every slab but the last has ntile_max tiles; the last may have fewer,
    // and its workgroups add into the first slots only.
    const int blocks = nstrip * cdiv(w, FitTile::BN);
    hipLaunchKernelGGL(k_fit_resid, dim3(blocks), dim3(FitTile::NT), 0, st, (int)ng, rk, d_theta,
                       (const double*)Cs, w, d_psi, ld, s1, s2, K2, q0, first, part);
    ISDF_HIP_LAUNCHED("k_fit_resid");
    if (first) nb_used = blocks;
    first = 0;
  }
  hipLaunchKernelGGL(k_fit_total, dim3(1), dim3(1024), 0, st, (const double*)part, nb_used,
                     d_out);
  ISDF_HIP_LAUNCHED("k_fit_total");
  double h[2] = {0.0, 0.0};
  ISDF_HIP_TRY(hipMemcpyAsync(h, d_out, sizeof(h), hipMemcpyDeviceToHost, st));
  ISDF_HIP_TRY(hipStreamSynchronize(st));
  out->fit_error = std::sqrt(h[0]) / std::sqrt(h[1]);
  out->z_norm = std::sqrt(h[1]);
  out->coulomb_error = -1.0;
  out->coulomb_norm = -1.0;
  return ISDF_OK;
}

int release_fit_buffer() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return ISDF_EHIP;
  if ((int)g_fit.size() > dev && g_fit[dev].ptr) {
    if (hipFree(g_fit[dev].ptr) != hipSuccess) return ISDF_EHIP;
    g_fit[dev].ptr = nullptr;
    g_fit[dev].bytes = 0;
  }
  return ISDF_OK;
}

}  // namespace isdf_hip

extern "C" int isdf_hip_fit_error(const isdf_problem* prob, const double* d_psi,
                                  const double* d_theta, const int* h_piv, void* stream,
                                  isdf_check* out) {
  std::lock_guard<std::mutex> lk(isdf_hip::solve_mutex());
  isdf_hip::set_error(0, "");
  const int rc =
      isdf_hip::fit_error_unlocked(prob, d_psi, d_theta, h_piv, (hipStream_t)stream, out);
  if (rc == ISDF_EHIP) (void)hipStreamSynchronize((hipStream_t)stream);
  return rc;
}
