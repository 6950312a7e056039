// LDS-staged FP64 MFMA GEMM building block for gfx950 (v_mfma_f64_16x16x4f64).
//
// One workgroup computes a BM x BN tile of A (BM x K) * B (K x BN), streaming K
// in chunks of KB through double-buffered LDS: the global loads of chunk c+1
// are in flight while the MFMAs of chunk c run.  Waves are arranged WGM x WGN;
// each wave owns TM x TN 16x16 accumulator tiles.
//
// Operand layouts (column-major matrices as LibISDF stores them):
//   A: element (m, k) at A[m + k*lda]                      (m contiguous)
//   B: B_KMAJOR: element (k, n) at B[k*ldb + n]            (n contiguous)
//      otherwise element (k, n) at B[k + n*ldb]            (k contiguous)
// Rows m >= mlim, columns n >= nlim and k >= K read as zero.
//
// MFMA lane layout (measured on gfx950): lane l supplies A[l%16][l/16] and
// B[l/16][l%16] and holds D[l/16 + 4r][l%16] in register r.
#pragma once

#include <hip/hip_runtime.h>

namespace isdf_hip {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <int BM_, int BN_, int KB_, int WGM_, int WGN_, bool B_KMAJOR_ = true>
struct MfmaTile {
  static constexpr int BM = BM_, BN = BN_, KB = KB_, WGM = WGM_, WGN = WGN_;
  static constexpr bool B_KMAJOR = B_KMAJOR_;
  static constexpr int NT = 64 * WGM * WGN;
  static constexpr int TM = BM / WGM / 16;
  static constexpr int TN = BN / WGN / 16;
  static constexpr int APT = BM * KB / NT;  // A elements each thread stages per chunk
  static constexpr int BPT = BN * KB / NT;
  static constexpr int LDS_DOUBLES = 2 * KB * (BM + BN);
  static_assert(TM >= 1 && TN >= 1, "tile too small for the wave grid");
  static_assert(BM * KB % NT == 0 && BN * KB % NT == 0, "chunk must split evenly");
  static_assert(KB % 4 == 0, "KB must be a multiple of 4");

  typedef d4 Acc[TM][TN];

  __device__ static void zero(Acc& acc) {
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  }

  // Global -> registers for the chunk starting at k0.  Element q of a thread
  // is tile element e = t + q*NT, so a wave touches 64 consecutive addresses.
  __device__ static void load(const double* __restrict__ A, long long lda,
                              const double* __restrict__ B, long long ldb, int K, int k0,
                              int mlim, int nlim, double (&ra)[APT], double (&rb)[BPT]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < APT; q++) {
      const int e = t + q * NT;
      const int kk = e / BM, m = e % BM;
      const int k = k0 + kk;
      ra[q] = (k < K && m < mlim) ? A[(long long)k * lda + m] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < BPT; q++) {
      const int e = t + q * NT;
      if constexpr (B_KMAJOR) {
        const int kk = e / BN, n = e % BN;
        const int k = k0 + kk;
        rb[q] = (k < K && n < nlim) ? B[(long long)k * ldb + n] : 0.0;
      } else {
        const int kk = e % KB, n = e / KB;
        const int k = k0 + kk;
        rb[q] = (k < K && n < nlim) ? B[(long long)n * ldb + k] : 0.0;
      }
    }
  }

  __device__ static void store(double* lds, int buf, const double (&ra)[APT],
                               const double (&rb)[BPT]) {
    double* As = lds + buf * KB * (BM + BN);
    double* Bs = As + KB * BM;
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < APT; q++) As[t + q * NT] = ra[q];  // As[kk][m]
#pragma unroll
    for (int q = 0; q < BPT; q++) {
      const int e = t + q * NT;
      if constexpr (B_KMAJOR) {
        Bs[e] = rb[q];  // Bs[kk][n]
      } else {
        Bs[(e % KB) * BN + e / KB] = rb[q];
      }
    }
  }

  __device__ static void compute(const double* lds, int buf, Acc& acc) {
    const double* As = lds + buf * KB * (BM + BN);
    const double* Bs = As + KB * BM;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % WGM, wn = wave / WGM;
    const int lr = lane & 15, lk = lane >> 4;
    const double* Aw = As + wm * (TM * 16) + lr;
    const double* Bw = Bs + wn * (TN * 16) + lr;
#pragma unroll
    for (int kk = 0; kk < KB; kk += 4) {
      double a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; i++) a[i] = Aw[(kk + lk) * BM + 16 * i];
#pragma unroll
      for (int j = 0; j < TN; j++) b[j] = Bw[(kk + lk) * BN + 16 * j];
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = mfma_f64(a[i], b[j], acc[i][j]);
    }
  }

  // acc += A[0:mlim, 0:K] * B[0:K, 0:nlim] over the whole K range.  Ends with a
  // workgroup barrier, so afterwards every thread's global reads are done.
  __device__ static void run(const double* __restrict__ A, long long lda,
                             const double* __restrict__ B, long long ldb, int K, int mlim,
                             int nlim, double* lds, Acc& acc) {
    double ra[APT], rb[BPT];
    const int nch = (K + KB - 1) / KB;
    if (nch == 0) return;
    load(A, lda, B, ldb, K, 0, mlim, nlim, ra, rb);
    store(lds, 0, ra, rb);
    __syncthreads();
    for (int c = 0; c < nch; c++) {
      const bool more = c + 1 < nch;
      if (more) load(A, lda, B, ldb, K, (c + 1) * KB, mlim, nlim, ra, rb);
      compute(lds, c & 1, acc);
      if (more) store(lds, (c + 1) & 1, ra, rb);
      __syncthreads();
    }
  }

  // Row / column of accumulator element (i, j, r) for this lane, relative to
  // the workgroup tile origin.
  __device__ static int row(int i, int r) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    return (wave % WGM) * (TM * 16) + 16 * i + (lane >> 4) + 4 * r;
  }
  __device__ static int col(int j) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    return (wave / WGM) * (TN * 16) + 16 * j + (lane & 15);
  }
};

}  // namespace isdf_hip
