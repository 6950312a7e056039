// Interpolation-basis construction for gfx950: Psi(mu) gather, the fused
// Z C^T = (Psi1 Psi1(mu)^T) o (Psi2 Psi2(mu)^T) MFMA kernel, C C^T extraction
// with the trace-scaled diagonal shift, the blocked FP64 Cholesky, and the
// blocked right-side triangular solves that turn Z C^T into Theta in place.
//
// Reference stages (LibISDF CPU path, the oracle): src/cpu/getbasis_scalapack.cpp
// :249-269 (gather), :297-369 (pdgemm + Hadamard), :390-428 (C C^T + eps),
// :455 (pdpotrf 'L'), :469-478 (pdtrsm R,L,T,N then R,L,N,N).
//
// FP64 matrix work runs on v_mfma_f64_16x16x4f64.  Operand fragments are read
// straight from global memory (they are re-used from the vector L1/L2): each
// lane of a 16x16x4 step supplies A[m = lane%16][k = lane/16] and
// B[k = lane/16][n = lane%16] and owns D[lane/16 + 4r][lane%16], r = 0..3
// (the gfx950 output layout, measured).

#include <hip/hip_runtime.h>

#include <algorithm>

#include "isdf_internal.h"

namespace isdf_hip {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

// psimuT[k*rk + mu] = psi(piv[mu], k): the gathered Psi(mu) stored k-major so
// that the GEMM's B operand is contiguous along mu.
__global__ __launch_bounds__(256) void k_gather(const double* __restrict__ psi, long long ld,
                                                int nstate, const int* __restrict__ piv, int rk,
                                                double* __restrict__ psimuT) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)rk * nstate) return;
  int mu = (int)(t % rk), k = (int)(t / rk);
  psimuT[t] = psi[(long long)k * ld + piv[mu]];
}

// acc[i][j] += A[m0+16i.., 0:K] * B[0:K, n0+16j..] with A column-major (ld lda)
// and B k-major (ld ldb); fragments double-buffered in registers.
template <int TM, int TN>
__device__ __forceinline__ void mfma_panel(const double* __restrict__ A, long long lda,
                                           const double* __restrict__ B, long long ldb, int K,
                                           long long m0, int n0, int lr, int lk, const bool (&mv)[TM],
                                           const bool (&nv)[TN], d4 (&acc)[TM][TN]) {
  double a[TM], b[TN], an[TM], bn[TN];
  {
    const bool kv = lk < K;
#pragma unroll
    for (int i = 0; i < TM; i++)
      a[i] = (kv && mv[i]) ? A[m0 + 16 * i + lr + (long long)lk * lda] : 0.0;
#pragma unroll
    for (int j = 0; j < TN; j++) b[j] = (kv && nv[j]) ? B[(long long)lk * ldb + n0 + 16 * j + lr] : 0.0;
  }
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int k = k0 + 4 + lk;
    const bool kv = k < K;
#pragma unroll
    for (int i = 0; i < TM; i++)
      an[i] = (kv && mv[i]) ? A[m0 + 16 * i + lr + (long long)k * lda] : 0.0;
#pragma unroll
    for (int j = 0; j < TN; j++)
      bn[j] = (kv && nv[j]) ? B[(long long)k * ldb + n0 + 16 * j + lr] : 0.0;
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) acc[i][j] = mfma_f64(a[i], b[j], acc[i][j]);
#pragma unroll
    for (int i = 0; i < TM; i++) a[i] = an[i];
#pragma unroll
    for (int j = 0; j < TN; j++) b[j] = bn[j];
  }
}

// C (ng x rk, column-major, ld ldc) = P1 o P1 or P1 o P2, with
// P1 = A1 (ng x K1, ld lda) * B1 (K1 x rk, k-major, ld ldb) and likewise P2.
// Workgroup: 4 waves in 2 x 2; a wave owns TM x TN 16x16 tiles.  The factors
// live only in accumulators: the epilogue writes their product once.
template <int TM, int TN, bool DUAL>
__global__ __launch_bounds__(256) void k_zct(int ng, int rk, const double* __restrict__ A1,
                                             const double* __restrict__ A2, long long lda, int K1,
                                             int K2, const double* __restrict__ B1,
                                             const double* __restrict__ B2, int ldb,
                                             double* __restrict__ C, long long ldc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const long long m0 = (long long)blockIdx.x * (32 * TM) + (wave & 1) * (16 * TM);
  const int n0 = blockIdx.y * (32 * TN) + (wave >> 1) * (16 * TN);

  d4 acc1[TM][TN];
  d4 acc2[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc1[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  if constexpr (DUAL) {
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) acc2[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  }

  bool mv[TM], nv[TN];
#pragma unroll
  for (int i = 0; i < TM; i++) mv[i] = (m0 + 16 * i + lr) < ng;
#pragma unroll
  for (int j = 0; j < TN; j++) nv[j] = (n0 + 16 * j + lr) < rk;

  mfma_panel<TM, TN>(A1, lda, B1, ldb, K1, m0, n0, lr, lk, mv, nv, acc1);
  if constexpr (DUAL) mfma_panel<TM, TN>(A2, lda, B2, ldb, K2, m0, n0, lr, lk, mv, nv, acc2);

#pragma unroll
  for (int i = 0; i < TM; i++) {
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const int nn = n0 + 16 * j + lr;
      if (!nv[j]) continue;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const long long m = m0 + 16 * i + lk + 4 * r;
        if (m < ng) {
          double v = acc1[i][j][r];
          double o;
          if constexpr (DUAL) o = v * acc2[i][j][r];
          else o = v * v;
          C[m + (long long)nn * ldc] = o;
        }
      }
    }
  }
}

// P(mu, nu) = Xi(piv[mu], nu), src/cpu/getbasis_scalapack.cpp:390-396.
__global__ __launch_bounds__(256) void k_cct_gather(const double* __restrict__ C, long long ldc,
                                                    const int* __restrict__ piv, int rk,
                                                    double* __restrict__ P) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)rk * rk) return;
  int mu = (int)(t % rk), nu = (int)(t / rk);
  P[t] = C[(long long)nu * ldc + piv[mu]];
}

// trace summed in ascending order, eps = 1e-10 * trace / rk added to the
// diagonal (getbasis_scalapack.cpp:398-428).  One block.
__global__ __launch_bounds__(256) void k_trace_eps(double* __restrict__ P, int rk,
                                                   double* __restrict__ eps_out) {
  __shared__ double eps_s;
  if (threadIdx.x == 0) {
    double tr = 0.0;
    for (int nu = 0; nu < rk; nu++) tr += P[(long long)nu * rk + nu];
    double eps = 1e-10 * tr / (double)rk;
    eps_s = eps;
    *eps_out = eps;
  }
  __syncthreads();
  const double eps = eps_s;
  for (int nu = threadIdx.x; nu < rk; nu += blockDim.x) P[(long long)nu * rk + nu] += eps;
}

// ---------------------------------------------------------------------------
// Blocked right-looking Cholesky, lower (the reference's pdpotrf 'L').
// ---------------------------------------------------------------------------

// Factor the jb x jb diagonal block at J0 (one wave, column-major in LDS), write
// L_JJ back, and its inverse into Dinv[J] (64 x 64, column-major, zero padded).
__global__ __launch_bounds__(64) void k_potrf_diag(double* __restrict__ A, int n, int J0,
                                                   double* __restrict__ Dinv,
                                                   int* __restrict__ info) {
  __shared__ double at[kNB][kNB];   // at[col][row]
  __shared__ double inv[kNB][kNB];  // inv[row][col]
  if (*info != 0) return;
  const int lane = threadIdx.x;
  const int jb = min(kNB, n - J0);
  for (int c = 0; c < kNB; c++) {
    double v = 0.0;
    if (c < jb && lane < jb && lane >= c) v = A[(long long)(J0 + c) * n + J0 + lane];
    at[c][lane] = v;
  }
  __syncthreads();
  int fail = 0;
  for (int j = 0; j < jb; j++) {
    double d = at[j][j];
    if (!(d > 0.0)) {
      fail = j + 1;
      break;
    }
    d = sqrt(d);
    __syncthreads();
    if (lane == j) at[j][j] = d;
    if (lane > j && lane < jb) at[j][lane] = at[j][lane] / d;
    __syncthreads();
    if (lane > j && lane < jb) {
      const double lij = at[j][lane];
      for (int k = j + 1; k <= lane; k++) at[k][lane] -= lij * at[j][k];
    }
    __syncthreads();
  }
  if (fail) {
    if (lane == 0) atomicCAS(info, 0, J0 + fail);
    return;
  }
  for (int c = 0; c < jb; c++)
    if (lane >= c && lane < jb) A[(long long)(J0 + c) * n + J0 + lane] = at[c][lane];
  // Column `lane` of inv(L_JJ) by forward substitution.
  const int c = lane;
  for (int i = 0; i < kNB; i++) {
    double val = 0.0;
    if (c < jb && i < jb) {
      if (i == c) {
        val = 1.0 / at[c][c];
      } else if (i > c) {
        double s = 0.0;
        for (int t = c; t < i; t++) s += at[t][i] * inv[t][c];
        val = -s / at[i][i];
      }
    }
    inv[i][c] = val;
    __syncthreads();
  }
  double* D = Dinv + (long long)(J0 / kNB) * kNB * kNB;
  for (int col = 0; col < kNB; col++) D[(long long)col * kNB + lane] = inv[lane][col];
}

// L_IJ = A_IJ inv(L_JJ)^T for the rows below the diagonal block; one wave per
// 64 rows.
__global__ __launch_bounds__(64) void k_potrf_panel(double* __restrict__ A, int n, int J0,
                                                    const double* __restrict__ Dinv,
                                                    const int* __restrict__ info) {
  __shared__ double at[kNB][kNB];  // at[k][row]
  __shared__ double di[kNB][kNB];  // di[col][k] = inv(L_JJ)[col][k]
  if (*info != 0) return;
  const int lane = threadIdx.x;
  const int jb = min(kNB, n - J0);
  const int r0 = J0 + jb + blockIdx.x * kNB;
  const int r = r0 + lane;
  const double* D = Dinv + (long long)(J0 / kNB) * kNB * kNB;
  for (int k = 0; k < kNB; k++) {
    at[k][lane] = (k < jb && r < n) ? A[(long long)(J0 + k) * n + r] : 0.0;
    di[lane][k] = D[(long long)k * kNB + lane];
  }
  __syncthreads();
  if (r >= n) return;
  for (int c = 0; c < jb; c++) {
    double s = 0.0;
    for (int k = 0; k <= c; k++) s += at[k][lane] * di[c][k];
    A[(long long)(J0 + c) * n + r] = s;
  }
}

// Trailing update A_IK -= L_IJ L_KJ^T on the lower 64 x 64 tiles (I >= K).
__global__ __launch_bounds__(256) void k_potrf_syrk(double* __restrict__ A, int n, int J0,
                                                    const int* __restrict__ info) {
  __shared__ double li[kNB][kNB + 1];  // li[k][row]
  __shared__ double lk[kNB][kNB + 1];
  if (*info != 0) return;
  const int ti = blockIdx.x, tk = blockIdx.y;
  if (tk > ti) return;
  const int jb = min(kNB, n - J0);
  const int T0 = J0 + jb;
  const int I0 = T0 + ti * kNB, K0 = T0 + tk * kNB;
  const int tid = threadIdx.x;
  for (int e = tid; e < kNB * kNB; e += 256) {
    int row = e % kNB, k = e / kNB;
    li[k][row] = (k < jb && I0 + row < n) ? A[(long long)(J0 + k) * n + I0 + row] : 0.0;
    lk[k][row] = (k < jb && K0 + row < n) ? A[(long long)(J0 + k) * n + K0 + row] : 0.0;
  }
  __syncthreads();
  const int tx = tid & 15, ty = tid >> 4;
  double s[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) s[a][b] = 0.0;
  for (int k = 0; k < jb; k++) {
    double x[4], y[4];
#pragma unroll
    for (int a = 0; a < 4; a++) x[a] = li[k][tx + 16 * a];
#pragma unroll
    for (int b = 0; b < 4; b++) y[b] = lk[k][ty + 16 * b];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) s[a][b] += x[a] * y[b];
  }
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int i = I0 + tx + 16 * a;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int k = K0 + ty + 16 * b;
      if (i < n && k < n && k <= i) A[(long long)k * n + i] -= s[a][b];
    }
  }
}

// Solve operands.  Forward (X L^T = B): X_J = [X_0..X_{J-1}, B_J] F_J^T with
// F_JQ = -inv(L_JJ) L_JQ (Q < J) and F_JJ = inv(L_JJ); stored Ft[k*n + j] = F[j][k].
// Backward (Theta L = X): Theta_J = [X_J, Theta_{J+1}..] G_J with
// G_PJ = -L_PJ inv(L_JJ) (P > J) and G_JJ = inv(L_JJ); stored Gt[k*n + j] = G[k][j].
__global__ __launch_bounds__(256) void k_make_fg(const double* __restrict__ L, int n,
                                                 const double* __restrict__ Dinv,
                                                 double* __restrict__ Ft, double* __restrict__ Gt) {
  const int P = blockIdx.x, Q = blockIdx.y;
  if (Q > P) return;
  const int P0 = P * kNB, Q0 = Q * kNB;
  const int pb = min(kNB, n - P0), qb = min(kNB, n - Q0);
  const double* DP = Dinv + (long long)P * kNB * kNB;
  const double* DQ = Dinv + (long long)Q * kNB * kNB;
  for (int e = threadIdx.x; e < kNB * kNB; e += 256) {
    const int i = e % kNB, jj = e / kNB;  // row in block P, column in block Q
    if (i >= pb || jj >= qb) continue;
    double f, g;
    if (P == Q) {
      f = DP[(long long)jj * kNB + i];
      g = f;
    } else {
      double sf = 0.0, sg = 0.0;
      for (int t = 0; t <= i; t++) sf += DP[(long long)t * kNB + i] * L[(long long)(Q0 + jj) * n + P0 + t];
      for (int t = jj; t < qb; t++) sg += L[(long long)(Q0 + t) * n + P0 + i] * DQ[(long long)jj * kNB + t];
      f = -sf;
      g = -sg;
    }
    Ft[(long long)(Q0 + jj) * n + P0 + i] = f;
    Gt[(long long)(P0 + i) * n + Q0 + jj] = g;
  }
}

// Theta = Z (L L^T)^{-1} in place, one wave per 32-row strip: the forward
// sweep writes X_J over Z_J, the backward sweep writes Theta_J over X_J.
// Each block column is one MFMA GEMM over the already-solved columns.
__global__ __launch_bounds__(64) void k_potrs(double* __restrict__ T, long long ng, long long ldt,
                                              int n, const double* __restrict__ Ft,
                                              const double* __restrict__ Gt) {
  constexpr int TM = 2, TN = kNB / 16;
  const int lane = threadIdx.x;
  const int lr = lane & 15, lk = lane >> 4;
  const long long m0 = (long long)blockIdx.x * (16 * TM);
  bool mv[TM];
#pragma unroll
  for (int i = 0; i < TM; i++) mv[i] = (m0 + 16 * i + lr) < ng;
  const int nJ = (n + kNB - 1) / kNB;

  for (int pass = 0; pass < 2; pass++) {
    const double* __restrict__ Bop = pass == 0 ? Ft : Gt;
    for (int s = 0; s < nJ; s++) {
      const int J = pass == 0 ? s : nJ - 1 - s;
      const int J0 = J * kNB;
      const int jb = min(kNB, n - J0);
      const int kbeg = pass == 0 ? 0 : J0;
      const int kend = pass == 0 ? J0 + jb : n;
      d4 acc[TM][TN];
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      bool nv[TN];
#pragma unroll
      for (int j = 0; j < TN; j++) nv[j] = (16 * j + lr) < jb;
      // Forward: K runs over columns [0, J0+jb); backward: over [J0, n).
      const double* Ap = T + (long long)kbeg * ldt;
      const double* Bp = Bop + (long long)kbeg * n + J0;
      mfma_panel<TM, TN>(Ap, ldt, Bp, n, kend - kbeg, m0, 0, lr, lk, mv, nv, acc);
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) {
          if (!nv[j]) continue;
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const long long m = m0 + 16 * i + lk + 4 * r;
            if (m < ng) T[m + (long long)(J0 + 16 * j + lr) * ldt] = acc[i][j][r];
          }
        }
      // The next block column reads what this one wrote (same wave, same rows).
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
  }
}

}  // namespace

int launch_gather_psimu(const double* psi, long long ld, int nstate, const int* piv, int rk,
                        double* psimuT, hipStream_t st) {
  hipLaunchKernelGGL(k_gather, dim3(cdiv((long long)rk * nstate, 256)), dim3(256), 0, st, psi, ld,
                     nstate, piv, rk, psimuT);
  ISDF_HIP_LAUNCHED("k_gather");
  return 0;
}

int launch_zct(const double* psi, long long ld, int ng, const double* psimuT, int rk, int s1,
               int K1, int s2, int K2, bool dual, double* C, long long ldc, hipStream_t st) {
  const double* A1 = psi + (long long)s1 * ld;
  const double* A2 = psi + (long long)s2 * ld;
  const double* B1 = psimuT + (long long)s1 * rk;
  const double* B2 = psimuT + (long long)s2 * rk;
  if (!dual) {
    constexpr int TM = 4, TN = 4;
    dim3 g(cdiv(ng, 32 * TM), cdiv(rk, 32 * TN));
    hipLaunchKernelGGL((k_zct<TM, TN, false>), g, dim3(256), 0, st, ng, rk, A1, A1, ld, K1, K1, B1,
                       B1, rk, C, ldc);
  } else {
    constexpr int TM = 4, TN = 2;
    dim3 g(cdiv(ng, 32 * TM), cdiv(rk, 32 * TN));
    hipLaunchKernelGGL((k_zct<TM, TN, true>), g, dim3(256), 0, st, ng, rk, A1, A2, ld, K1, K2, B1,
                       B2, rk, C, ldc);
  }
  ISDF_HIP_LAUNCHED("k_zct");
  return 0;
}

int launch_cct(const double* C, long long ldc, const int* piv, int rk, double* P, double* eps_dev,
               hipStream_t st) {
  hipLaunchKernelGGL(k_cct_gather, dim3(cdiv((long long)rk * rk, 256)), dim3(256), 0, st, C, ldc,
                     piv, rk, P);
  ISDF_HIP_LAUNCHED("k_cct_gather");
  hipLaunchKernelGGL(k_trace_eps, dim3(1), dim3(256), 0, st, P, rk, eps_dev);
  ISDF_HIP_LAUNCHED("k_trace_eps");
  return 0;
}

int launch_potrf(double* P, int rk, double* Dinv, int* info_dev, hipStream_t st) {
  const int nJ = (rk + kNB - 1) / kNB;
  for (int J = 0; J < nJ; J++) {
    const int J0 = J * kNB;
    const int jb = std::min(kNB, rk - J0);
    hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(64), 0, st, P, rk, J0, Dinv, info_dev);
    ISDF_HIP_LAUNCHED("k_potrf_diag");
    const int rest = rk - J0 - jb;
    if (rest > 0) {
      hipLaunchKernelGGL(k_potrf_panel, dim3(cdiv(rest, kNB)), dim3(64), 0, st, P, rk, J0, Dinv,
                         info_dev);
      ISDF_HIP_LAUNCHED("k_potrf_panel");
      const unsigned t = cdiv(rest, kNB);
      hipLaunchKernelGGL(k_potrf_syrk, dim3(t, t), dim3(256), 0, st, P, rk, J0, info_dev);
      ISDF_HIP_LAUNCHED("k_potrf_syrk");
    }
  }
  return 0;
}

int launch_make_fg(const double* L, int rk, const double* Dinv, double* Ft, double* Gt,
                   hipStream_t st) {
  const unsigned nJ = cdiv(rk, kNB);
  hipLaunchKernelGGL(k_make_fg, dim3(nJ, nJ), dim3(256), 0, st, L, rk, Dinv, Ft, Gt);
  ISDF_HIP_LAUNCHED("k_make_fg");
  return 0;
}

int launch_potrs(double* T, long long ng, long long ldt, int rk, const double* Ft, const double* Gt,
                 hipStream_t st) {
  hipLaunchKernelGGL(k_potrs, dim3(cdiv(ng, 32)), dim3(64), 0, st, T, ng, ldt, rk, Ft, Gt);
  ISDF_HIP_LAUNCHED("k_potrs");
  return 0;
}

}  // namespace isdf_hip
