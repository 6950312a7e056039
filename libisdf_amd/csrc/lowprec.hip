// Mixed-precision basis construction for gfx950.
//
// The FP64 path (basis.hip) forms Theta = Z C^T (L L^T)^{-1} with FP64 MFMA.
// The mixed forms keep K-means, the Cholesky and the inverse in FP64 and move
// the two O(Ng) contractions to low-precision MFMA with FP32 accumulation:
//
//   Z C^T   Zt[g][nu]  = (sum_k Psi1(g,k) Psi1mu(nu,k)) * (sum_k Psi2(g,k) Psi2mu(nu,k))
//   Theta   Theta(g,nu) = sum_k Zt[g][k] W(k,nu),  W = (L L^T)^{-1} (FP64, rounded once)
//
// Two arithmetic modes share one tile (LpTile):
//   FP32    v_mfma_f32_16x16x4f32 on FP32 operands;
//   BF16X3  each FP32 operand x is split exactly into three bf16 parts
//           x = h + m + l (round-to-nearest each, 27 bits cover the 24-bit
//           significand) and the six products hh, hm, mh, hl, mm, lh run on
//           v_mfma_f32_16x16x32_bf16; the three dropped ones are below 2^-26.
//
// Both contractions are "NT" GEMMs, C[m][n] = sum_k A[m][k] B[n][k], with both
// operands k-contiguous in HBM (FP32, k padded to 32 with zeros).  That is the
// order the MFMA operands want, so global->LDS staging is 16-byte vector
// traffic and Z C^T is stored g-major (Zt), which is the B operand of Theta.
// Theta is produced transposed (rows nu, columns g) so that each lane group
// writes 16 consecutive grid points of one FP64 column of the caller's
// column-major theta.
//
// Reference stages replaced (src/cpu/getbasis_scalapack.cpp): :263-269 gather,
// :297-369 pdgemm + Hadamard, :390-428 C C^T + eps, :469-478 the two pdtrsm.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "isdf_internal.h"
#include "mfma_gemm.h"

namespace isdf_hip {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

// x = h + m + l exactly: each part is the bf16 rounding of what is left.
__device__ __forceinline__ void split3(f4 x, bf4& h, bf4& m, bf4& l) {
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const __bf16 a = (__bf16)x[e];
    const float r1 = x[e] - (float)a;
    const __bf16 b = (__bf16)r1;
    const float r2 = r1 - (float)b;
    h[e] = a;
    m[e] = b;
    l[e] = (__bf16)r2;
  }
}

// C[m][n] = sum_k A[m][k] B[n][k] over a BM x BN workgroup tile, K streamed in
// chunks of 32 through double-buffered LDS.  A(m, k) = A[m*lda + k],
// B(n, k) = B[n*ldb + k]; K, lda and ldb are multiples of 4 (K of 32) and the
// operands are 16-byte aligned.  Rows m >= mlim and n >= nlim read as zero.
//
// 16x16 MFMA output layout (FP32 accumulators, measured on gfx950): lane l
// holds D[4*(l/16) + r][l % 16] in register r.
template <int BM_, int BN_, int WGM_, int WGN_, bool BF16_>
struct LpTile {
  static constexpr int BM = BM_, BN = BN_, WGM = WGM_, WGN = WGN_;
  static constexpr bool BF16 = BF16_;
  static constexpr int NT = 64 * WGM * WGN;
  static constexpr int TM = BM / WGM / 16;
  static constexpr int TN = BN / WGN / 16;
  static constexpr int KB = 32;
  static constexpr int APT = BM * (KB / 4) / NT;  // float4 per thread per chunk
  static constexpr int BPT = BN * (KB / 4) / NT;
  static constexpr int RS32 = KB + 4;  // FP32 row stride in LDS (floats): no bank conflicts
  static constexpr int RS16 = KB + 8;  // bf16 plane row stride (elements), 80 bytes
  static constexpr int BUF_BYTES = BF16 ? 3 * (BM + BN) * RS16 * 2 : (BM + BN) * RS32 * 4;
  static constexpr int LDS_BYTES = 2 * BUF_BYTES;
  static_assert(TM >= 1 && TN >= 1, "tile too small for the wave grid");
  static_assert(APT >= 1 && BPT >= 1 && BM * 8 % NT == 0 && BN * 8 % NT == 0, "staging split");
  static_assert(LDS_BYTES <= 160 * 1024, "LDS");

  typedef f4 Acc[TM][TN];

  __device__ static void zero(Acc& acc) {
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  }

  // Element q of a thread is float4 e = t + q*NT of the chunk: row e/8,
  // k offset 4*(e%8); eight consecutive threads read one 128-byte row piece.
  __device__ static void load(const float* __restrict__ A, long long lda,
                              const float* __restrict__ B, long long ldb, int K, int k0, int mlim,
                              int nlim, f4 (&ra)[APT], f4 (&rb)[BPT]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < APT; q++) {
      const int e = t + q * NT;
      const int m = e >> 3, k = k0 + 4 * (e & 7);
      ra[q] = (m < mlim && k < K) ? *(const f4*)(A + (long long)m * lda + k)
                                  : f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int q = 0; q < BPT; q++) {
      const int e = t + q * NT;
      const int n = e >> 3, k = k0 + 4 * (e & 7);
      rb[q] = (n < nlim && k < K) ? *(const f4*)(B + (long long)n * ldb + k)
                                  : f4{0.f, 0.f, 0.f, 0.f};
    }
  }

  __device__ static void store(char* lds, int buf, const f4 (&ra)[APT], const f4 (&rb)[BPT]) {
    const int t = threadIdx.x;
    if constexpr (!BF16) {
      float* As = (float*)(lds + buf * BUF_BYTES);
      float* Bs = As + BM * RS32;
#pragma unroll
      for (int q = 0; q < APT; q++) {
        const int e = t + q * NT;
        *(f4*)(As + (e >> 3) * RS32 + 4 * (e & 7)) = ra[q];
      }
#pragma unroll
      for (int q = 0; q < BPT; q++) {
        const int e = t + q * NT;
        *(f4*)(Bs + (e >> 3) * RS32 + 4 * (e & 7)) = rb[q];
      }
    } else {
      __bf16* Ap = (__bf16*)(lds + buf * BUF_BYTES);  // planes h, m, l of A
      __bf16* Bp = Ap + 3 * BM * RS16;                // planes h, m, l of B
#pragma unroll
      for (int q = 0; q < APT; q++) {
        const int e = t + q * NT;
        const int o = (e >> 3) * RS16 + 4 * (e & 7);
        bf4 h, m, l;
        split3(ra[q], h, m, l);
        *(bf4*)(Ap + o) = h;
        *(bf4*)(Ap + BM * RS16 + o) = m;
        *(bf4*)(Ap + 2 * BM * RS16 + o) = l;
      }
#pragma unroll
      for (int q = 0; q < BPT; q++) {
        const int e = t + q * NT;
        const int o = (e >> 3) * RS16 + 4 * (e & 7);
        bf4 h, m, l;
        split3(rb[q], h, m, l);
        *(bf4*)(Bp + o) = h;
        *(bf4*)(Bp + BN * RS16 + o) = m;
        *(bf4*)(Bp + 2 * BN * RS16 + o) = l;
      }
    }
  }

  __device__ static void compute(const char* lds, int buf, Acc& acc) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % WGM, wn = wave / WGM;
    const int lr = lane & 15, lk = lane >> 4;
    if constexpr (!BF16) {
      const float* As = (const float*)(lds + buf * BUF_BYTES);
      const float* Bs = As + BM * RS32;
      const float* Aw = As + (wm * TM * 16 + lr) * RS32 + lk;
      const float* Bw = Bs + (wn * TN * 16 + lr) * RS32 + lk;
#pragma unroll
      for (int kk = 0; kk < KB; kk += 4) {
        float a[TM], b[TN];
#pragma unroll
        for (int i = 0; i < TM; i++) a[i] = Aw[16 * i * RS32 + kk];
#pragma unroll
        for (int j = 0; j < TN; j++) b[j] = Bw[16 * j * RS32 + kk];
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
      }
    } else {
      const __bf16* Ap = (const __bf16*)(lds + buf * BUF_BYTES);
      const __bf16* Bp = Ap + 3 * BM * RS16;
      const __bf16* Aw = Ap + (wm * TM * 16 + lr) * RS16 + 8 * lk;
      const __bf16* Bw = Bp + (wn * TN * 16 + lr) * RS16 + 8 * lk;
      bf8 ah[TM], am[TM], al[TM];
#pragma unroll
      for (int i = 0; i < TM; i++) {
        ah[i] = *(const bf8*)(Aw + 16 * i * RS16);
        am[i] = *(const bf8*)(Aw + BM * RS16 + 16 * i * RS16);
        al[i] = *(const bf8*)(Aw + 2 * BM * RS16 + 16 * i * RS16);
      }
#pragma unroll
      for (int j = 0; j < TN; j++) {
        const bf8 bh = *(const bf8*)(Bw + 16 * j * RS16);
        const bf8 bm = *(const bf8*)(Bw + BN * RS16 + 16 * j * RS16);
        const bf8 bl = *(const bf8*)(Bw + 2 * BN * RS16 + 16 * j * RS16);
#pragma unroll
        for (int i = 0; i < TM; i++) {
          // smallest terms first
          f4 c = acc[i][j];
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[i], bm, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bm, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[i], bh, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh, c, 0, 0, 0);
          acc[i][j] = c;
        }
      }
    }
  }

  // acc += A[0:mlim, 0:K] B[0:nlim, 0:K]^T.  Ends with a workgroup barrier.
  __device__ static void run(const float* __restrict__ A, long long lda,
                             const float* __restrict__ B, long long ldb, int K, int mlim,
                             int nlim, char* lds, Acc& acc) {
    f4 ra[APT], rb[BPT];
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

  __device__ static int row(int i, int r) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    return (wave % WGM) * (TM * 16) + 16 * i + 4 * (lane >> 4) + r;
  }
  __device__ static int col(int j) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    return (wave / WGM) * (TN * 16) + 16 * j + (lane & 15);
  }
};

// Workgroup b -> (strip, tile): the ntile tiles of one strip of the shared
// operand land on one XCD (workgroups are dealt round-robin to the 8 XCDs) and
// run back to back, so the strip is read from HBM once into that XCD's L2.
__device__ __forceinline__ void xcd_map(int b, int ntile, int nstrip, int* s, int* t) {
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

// Packs columns [c0, c0+K) of the FP64 column-major psi (ld) into FP32 rows
// out[g*kp + k], zero for K <= k < kp.  64 grid points x 32 states per block,
// transposed through LDS so that both sides are coalesced.
__global__ __launch_bounds__(256) void k_pack_rows(const double* __restrict__ psi, long long ld,
                                                   int ng, int c0, int K, int kp,
                                                   float* __restrict__ out) {
  __shared__ float tile[32][65];
  const long long g0 = (long long)blockIdx.x * 64;
  const int k0 = blockIdx.y * 32;
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int kk = (t >> 6) + 4 * q, gg = t & 63;
    const long long g = g0 + gg;
    const int k = k0 + kk;
    tile[kk][gg] = (g < ng && k < K) ? (float)psi[(long long)(c0 + k) * ld + g] : 0.f;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int gg = (t >> 5) + 8 * q, kk = t & 31;
    const long long g = g0 + gg;
    if (g < ng) out[g * kp + k0 + kk] = tile[kk][gg];
  }
}

// out[nu*kp + k] = psi(piv[nu], c0 + k) in FP32, zero for k >= K.
__global__ __launch_bounds__(256) void k_gather_rows(const double* __restrict__ psi, long long ld,
                                                     const int* __restrict__ piv, int rk, int c0,
                                                     int K, int kp, float* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)rk * kp) return;
  const int k = (int)(t % kp), nu = (int)(t / kp);
  out[t] = k < K ? (float)psi[(long long)(c0 + k) * ld + piv[nu]] : 0.f;
}

// Zt[g][nu] = P1(g, nu) * P2(g, nu) (P2 = P1 when the pair sets agree), for
// nu < ldz (columns rk..ldz come out zero, as the Theta GEMM's k padding).
template <class Tile, bool DUAL>
__global__ __launch_bounds__(Tile::NT, 1) void k_zct_lp(int ng, int rk, int ldz,
                                                        const float* __restrict__ A1,
                                                        const float* __restrict__ A2, int kp1,
                                                        int kp2, const float* __restrict__ B1,
                                                        const float* __restrict__ B2,
                                                        float* __restrict__ Zt) {
  __shared__ __attribute__((aligned(16))) char lds[Tile::LDS_BYTES];
  const int nstrip = (ng + Tile::BM - 1) / Tile::BM;
  const int ntile = (ldz + Tile::BN - 1) / Tile::BN;
  int s, ct;
  xcd_map(blockIdx.x, ntile, nstrip, &s, &ct);
  if (s >= nstrip) return;
  const long long m0 = (long long)s * Tile::BM;
  const int n0 = ct * Tile::BN;
  const int mlim = (int)std::min<long long>(Tile::BM, ng - m0);
  const int nlim = std::min(Tile::BN, rk - n0);   // operand rows that exist
  const int wlim = std::min(Tile::BN, ldz - n0);  // columns written (zeros past rk)
  typename Tile::Acc acc1, acc2;
  Tile::zero(acc1);
  Tile::run(A1 + m0 * kp1, kp1, B1 + (long long)n0 * kp1, kp1, kp1, mlim, nlim, lds, acc1);
  if constexpr (DUAL) {
    Tile::zero(acc2);
    Tile::run(A2 + m0 * kp2, kp2, B2 + (long long)n0 * kp2, kp2, kp2, mlim, nlim, lds, acc2);
  }
  float* Zm = Zt + m0 * ldz + n0;
#pragma unroll
  for (int j = 0; j < Tile::TN; j++) {
    const int n = Tile::col(j);
#pragma unroll
    for (int i = 0; i < Tile::TM; i++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = Tile::row(i, r);
        if (m < mlim && n < wlim) {
          const float v1 = acc1[i][j][r];
          const float v2 = DUAL ? acc2[i][j][r] : v1;
          Zm[(long long)m * ldz + n] = v1 * v2;
        }
      }
  }
}

// P(mu, nu) = Zt[piv[mu]][nu] (column-major rk x rk FP64),
// src/cpu/getbasis_scalapack.cpp:390-396.
__global__ __launch_bounds__(256) void k_cct_lp(const float* __restrict__ Zt, int ldz,
                                                const int* __restrict__ piv, int rk,
                                                double* __restrict__ P) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)rk * rk) return;
  const int mu = (int)(t % rk), nu = (int)(t / rk);
  P[t] = (double)Zt[(long long)piv[mu] * ldz + nu];
}

// T = I (n x n, column-major FP64).
__global__ __launch_bounds__(256) void k_identity(double* __restrict__ T, int n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * n) return;
  T[t] = (t % n == t / n) ? 1.0 : 0.0;
}

// Wf[nu*ldw + k] = W(k, nu) in FP32, zero for k >= rk (W column-major FP64).
__global__ __launch_bounds__(256) void k_pack_w(const double* __restrict__ W, int rk, int ldw,
                                                float* __restrict__ Wf) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)rk * ldw) return;
  const int k = (int)(t % ldw), nu = (int)(t / ldw);
  Wf[t] = k < rk ? (float)W[(long long)nu * rk + k] : 0.f;
}

// theta(g, nu) (=|+=) sum_k Wf[nu][k] Zt[g][k]: rows nu of the tile are the A
// operand, grid points g the B operand; each lane group writes 16 consecutive
// g of one column of the column-major theta.
template <class Tile, bool ADD>
__global__ __launch_bounds__(Tile::NT, 1) void k_theta_lp(int ng, int rk, int ldk,
                                                          const float* __restrict__ Wf,
                                                          const float* __restrict__ Zt,
                                                          double* __restrict__ theta,
                                                          long long ldt) {
  __shared__ __attribute__((aligned(16))) char lds[Tile::LDS_BYTES];
  const int nstrip = (ng + Tile::BN - 1) / Tile::BN;  // grid-point strips (shared Zt rows)
  const int ntile = (rk + Tile::BM - 1) / Tile::BM;   // nu tiles
  int s, vt;
  xcd_map(blockIdx.x, ntile, nstrip, &s, &vt);
  if (s >= nstrip) return;
  const long long g0 = (long long)s * Tile::BN;
  const int v0 = vt * Tile::BM;
  const int mlim = std::min(Tile::BM, rk - v0);
  const int nlim = (int)std::min<long long>(Tile::BN, ng - g0);
  typename Tile::Acc acc;
  Tile::zero(acc);
  Tile::run(Wf + (long long)v0 * ldk, ldk, Zt + g0 * ldk, ldk, ldk, mlim, nlim, lds, acc);
  double* Tt = theta + (long long)v0 * ldt + g0;
#pragma unroll
  for (int i = 0; i < Tile::TM; i++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int m = Tile::row(i, r);
#pragma unroll
      for (int j = 0; j < Tile::TN; j++) {
        const int n = Tile::col(j);
        if (m < mlim && n < nlim) {
          double* p = Tt + (long long)m * ldt + n;
          const double v = (double)acc[i][j][r];
          if constexpr (ADD) *p += v;
          else *p = v;
        }
      }
    }
}

// Residual of the FP64 correction step: Zt[g][nu] <- Zt[g][nu] - (Theta Pe)(g, nu)
// with Theta FP64 (column-major, ldt) and Pe = C C^T + eps I (column-major).
typedef MfmaTile<256, 128, 16, 4, 2, false> ResidTile;
__global__ __launch_bounds__(ResidTile::NT, 1) void k_resid(int ng, int rk,
                                                            const double* __restrict__ theta,
                                                            long long ldt,
                                                            const double* __restrict__ Pe,
                                                            float* __restrict__ Zt, int ldz) {
  __shared__ double lds[ResidTile::LDS_DOUBLES];
  const int nstrip = (ng + ResidTile::BM - 1) / ResidTile::BM;
  const int ntile = (rk + ResidTile::BN - 1) / ResidTile::BN;
  int s, ct;
  xcd_map(blockIdx.x, ntile, nstrip, &s, &ct);
  if (s >= nstrip) return;
  const long long m0 = (long long)s * ResidTile::BM;
  const int n0 = ct * ResidTile::BN;
  const int mlim = (int)std::min<long long>(ResidTile::BM, ng - m0);
  const int nlim = std::min(ResidTile::BN, rk - n0);
  typename ResidTile::Acc acc;
  ResidTile::zero(acc);
  // B(k, n) = Pe(k, n0+n) at Pe[(n0+n)*rk + k]: k contiguous.
  ResidTile::run(theta + m0, ldt, Pe + (long long)n0 * rk, rk, rk, mlim, nlim, lds, acc);
  ResidTile::epilogue(acc, [&](int m, int n, double v) {
    if (m < mlim && n < nlim) {
      float* z = Zt + (m0 + m) * (long long)ldz + n0 + n;
      *z = (float)((double)*z - v);
    }
  });
}

// Tile variants.  Both contractions use the same shapes per mode.
typedef LpTile<256, 128, 4, 2, false> LpF32_0;   // 8 waves of 64 x 64
typedef LpTile<128, 128, 2, 2, false> LpF32_1;   // 4 waves of 64 x 64
typedef LpTile<256, 256, 4, 2, false> LpF32_2;   // 8 waves of 64 x 128
typedef LpTile<128, 128, 2, 2, true> LpB16_0;    // 4 waves of 64 x 64
typedef LpTile<192, 128, 2, 2, true> LpB16_1;    // 4 waves of 96 x 64
typedef LpTile<128, 64, 2, 2, true> LpB16_2;     // 4 waves of 64 x 32

}  // namespace

int lp_variants(int mode) { return 3; }

int launch_pack_psi(const double* psi, long long ld, int ng, int c0, int K, int kp, float* out,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_pack_rows, dim3(cdiv(ng, 64), cdiv(kp, 32)), dim3(256), 0, st, psi, ld, ng,
                     c0, K, kp, out);
  ISDF_HIP_LAUNCHED("k_pack_rows");
  return 0;
}

int launch_gather_rows(const double* psi, long long ld, const int* piv, int rk, int c0, int K,
                       int kp, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_rows, dim3(cdiv((long long)rk * kp, 256)), dim3(256), 0, st, psi,
                     ld, piv, rk, c0, K, kp, out);
  ISDF_HIP_LAUNCHED("k_gather_rows");
  return 0;
}

int launch_zct_lp(int mode, int ng, int rk, int ldz, const float* A1, const float* A2, int kp1,
                  int kp2, const float* B1, const float* B2, bool dual, float* Zt, int variant,
                  hipStream_t st) {
  if (variant < 0) variant = 0;
  // Two 256 x 256 FP32 accumulators do not fit the register file.
  if (dual && mode == 1 && variant == 2) variant = 0;
#define ISDF_ZLP(TL)                                                                           \
  do {                                                                                         \
    const unsigned nb = cdiv(ng, TL::BM) * cdiv(ldz, TL::BN);                                  \
    if (dual)                                                                                  \
      hipLaunchKernelGGL((k_zct_lp<TL, true>), dim3(nb), dim3(TL::NT), 0, st, ng, rk, ldz, A1, \
                         A2, kp1, kp2, B1, B2, Zt);                                            \
    else                                                                                       \
      hipLaunchKernelGGL((k_zct_lp<TL, false>), dim3(nb), dim3(TL::NT), 0, st, ng, rk, ldz,    \
                         A1, A2, kp1, kp2, B1, B2, Zt);                                        \
  } while (0)
  if (mode == 1) {
    if (variant == 1) ISDF_ZLP(LpF32_1);
    else if (variant == 2) ISDF_ZLP(LpF32_2);
    else ISDF_ZLP(LpF32_0);
  } else {
    if (variant == 1) ISDF_ZLP(LpB16_1);
    else if (variant == 2) ISDF_ZLP(LpB16_2);
    else ISDF_ZLP(LpB16_0);
  }
#undef ISDF_ZLP
  ISDF_HIP_LAUNCHED("k_zct_lp");
  return 0;
}

int launch_cct_lp(const float* Zt, int ldz, const int* piv, int rk, double* P, double* eps_dev,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_cct_lp, dim3(cdiv((long long)rk * rk, 256)), dim3(256), 0, st, Zt, ldz,
                     piv, rk, P);
  ISDF_HIP_LAUNCHED("k_cct_lp");
  return launch_trace_eps(P, rk, eps_dev, st);
}

int launch_inverse(int rk, const double* Ft, const double* Gt, double* W, int ldw, float* Wf,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_identity, dim3(cdiv((long long)rk * rk, 256)), dim3(256), 0, st, W, rk);
  ISDF_HIP_LAUNCHED("k_identity");
  // W = I (L L^T)^{-1}: the blocked in-place solve with its 32-row tile (rk/32 workgroups).
  if (int rc = launch_potrs(W, rk, rk, rk, Ft, Gt, 4, st)) return rc;
  hipLaunchKernelGGL(k_pack_w, dim3(cdiv((long long)rk * ldw, 256)), dim3(256), 0, st, W, rk, ldw,
                     Wf);
  ISDF_HIP_LAUNCHED("k_pack_w");
  return 0;
}

int launch_theta_lp(int mode, int ng, int rk, int ldk, const float* Wf, const float* Zt,
                    double* theta, long long ldt, bool add, int variant, hipStream_t st) {
  if (variant < 0) variant = 0;
#define ISDF_TLP(TL)                                                                            \
  do {                                                                                          \
    const unsigned nb = cdiv(ng, TL::BN) * cdiv(rk, TL::BM);                                    \
    if (add)                                                                                    \
      hipLaunchKernelGGL((k_theta_lp<TL, true>), dim3(nb), dim3(TL::NT), 0, st, ng, rk, ldk, Wf, \
                         Zt, theta, ldt);                                                       \
    else                                                                                        \
      hipLaunchKernelGGL((k_theta_lp<TL, false>), dim3(nb), dim3(TL::NT), 0, st, ng, rk, ldk,   \
                         Wf, Zt, theta, ldt);                                                   \
  } while (0)
  if (mode == 1) {
    if (variant == 1) ISDF_TLP(LpF32_1);
    else if (variant == 2) ISDF_TLP(LpF32_2);
    else ISDF_TLP(LpF32_0);
  } else {
    if (variant == 1) ISDF_TLP(LpB16_1);
    else if (variant == 2) ISDF_TLP(LpB16_2);
    else ISDF_TLP(LpB16_0);
  }
#undef ISDF_TLP
  ISDF_HIP_LAUNCHED("k_theta_lp");
  return 0;
}

int launch_resid(int ng, int rk, const double* theta, long long ldt, const double* Pe, float* Zt,
                 int ldz, hipStream_t st) {
  const unsigned nb = cdiv(ng, ResidTile::BM) * cdiv(rk, ResidTile::BN);
  hipLaunchKernelGGL(k_resid, dim3(nb), dim3(ResidTile::NT), 0, st, ng, rk, theta, ldt, Pe, Zt,
                     ldz);
  ISDF_HIP_LAUNCHED("k_resid");
  return 0;
}

}  // namespace isdf_hip
