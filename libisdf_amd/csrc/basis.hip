// Interpolation-basis construction for gfx950: Psi(mu) gather, the fused
// Z C^T = (Psi1 Psi1(mu)^T) o (Psi2 Psi2(mu)^T) MFMA kernel, C C^T extraction
// with the trace-scaled diagonal shift, the blocked FP64 Cholesky, and the
// blocked right-side triangular solves that turn Z C^T into Theta in place.
//
// Reference stages (LibISDF CPU path, the oracle): src/cpu/getbasis_scalapack.cpp
// :249-269 (gather), :297-369 (pdgemm + Hadamard), :390-428 (C C^T + eps),
// :455 (pdpotrf 'L'), :469-478 (pdtrsm R,L,T,N then R,L,N,N).
//
// FP64 matrix work runs on v_mfma_f64_16x16x4f64 through the LDS-staged tile
// of mfma_gemm.h.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "isdf_internal.h"
#include "mfma_gemm.h"

namespace isdf_hip {

namespace {

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

// Tile shapes: 256 x 128 per workgroup (8 waves of 64 x 64) for the single
// product; 256 x 64 with two accumulators when the pair sets differ.
typedef MfmaTile<256, 128, 16, 4, 2> ZctTile0;      // 8 waves of 64 x 64
typedef MfmaTile<256, 128, 16, 4, 1> ZctTile1;      // 4 waves of 64 x 128
typedef ValuTile<256, 128, 16, 1, 8> ZctTile2;      // vector FMA, 8 waves of 256 x 16
typedef MfmaTile<256, 64, 16, 4, 2> ZctTileDual0;   // two accumulators
typedef MfmaTile<128, 64, 16, 2, 2> ZctTileDual1;

// Maps a linear workgroup id to (row strip, column tile) so that the ncolt
// column tiles of one row strip land on the same XCD (workgroups are dealt
// round-robin to the 8 XCDs) and run back to back, sharing the strip in L2.
// Row strips beyond the last full group of 8 are mapped in plain order.
__device__ __forceinline__ void xcd_tile(int b, int ncolt, int nrows, int* rs, int* ct) {
  const int full = (nrows / 8) * 8;
  if (b < full * ncolt) {
    const int x = b & 7, l = b >> 3;
    *rs = x + 8 * (l / ncolt);
    *ct = l % ncolt;
  } else {
    const int b2 = b - full * ncolt;
    *rs = full + b2 / ncolt;
    *ct = b2 % ncolt;
  }
}

// C (ng x rk, column-major, ld ldc) = P1 o P1 or P1 o P2 with
// P1 = A1 (ng x K1, ld lda) * B1 (K1 x rk, k-major, ld ldb), P2 likewise.  The
// factors live only in accumulators; the epilogue writes their product once.
template <class Tile, bool DUAL>
__global__ __launch_bounds__(Tile::NT, 1) void k_zct(int ng, int rk, const double* __restrict__ A1,
                                                const double* __restrict__ A2, long long lda,
                                                int K1, int K2, const double* __restrict__ B1,
                                                const double* __restrict__ B2, int ldb,
                                                double* __restrict__ C, long long ldc) {
  __shared__ double lds[Tile::LDS_DOUBLES];
  const int nrows = (ng + Tile::BM - 1) / Tile::BM;
  const int ncolt = (rk + Tile::BN - 1) / Tile::BN;
  int rs, ct;
  xcd_tile(blockIdx.x, ncolt, nrows, &rs, &ct);
  if (rs >= nrows) return;
  const long long m0 = (long long)rs * Tile::BM;
  const int n0 = ct * Tile::BN;
  const int mlim = (int)std::min<long long>(Tile::BM, ng - m0);
  const int nlim = std::min(Tile::BN, rk - n0);
  typename Tile::Acc acc1, acc2;
  Tile::zero(acc1);
  Tile::run(A1 + m0, lda, B1 + n0, ldb, K1, mlim, nlim, lds, acc1);
  if constexpr (DUAL) {
    Tile::zero(acc2);
    Tile::run(A2 + m0, lda, B2 + n0, ldb, K2, mlim, nlim, lds, acc2);
  }
  if constexpr (DUAL) {
    Tile::epilogue2(acc1, acc2, [&](int m, int n, double v1, double v2) {
      if (m < mlim && n < nlim) C[(long long)(n0 + n) * ldc + m0 + m] = v1 * v2;
    });
  } else {
    Tile::epilogue(acc1, [&](int m, int n, double v) {
      if (m < mlim && n < nlim) C[(long long)(n0 + n) * ldc + m0 + m] = v * v;
    });
  }
}

// k_zct with the epilogue staged through LDS: each half of the tile's columns
// (the columns of one wave column) is formed (P1 P1 or P1 P2) into LDS
// column-major, then written back as contiguous column segments (a wave stores
// 512 consecutive bytes) instead of the accumulator layout's 32-byte pieces.
template <class Tile, bool DUAL, bool PERSIST = false>
__global__ __launch_bounds__(Tile::NT, 1) void k_zct_lds_epilogue(
    int ng, int rk, const double* __restrict__ A1, const double* __restrict__ A2,
    long long lda, int K1, int K2, const double* __restrict__ B1,
    const double* __restrict__ B2, int ldb, double* __restrict__ C, long long ldc) {
  constexpr int HALF = Tile::BN / 2, CS = Tile::BM + 4;  // staged columns, column stride
  constexpr int STAGE = HALF * CS;
  static_assert(Tile::WGN == 2, "one wave column per half");
  __shared__ double lds[Tile::LDS_DOUBLES > STAGE ? Tile::LDS_DOUBLES : STAGE];
  const int nrows = (ng + Tile::BM - 1) / Tile::BM;
  const int ncolt = (rk + Tile::BN - 1) / Tile::BN;
  // One tile per workgroup, or (PERSIST: a smaller grid, a multiple of 8, so
  // tile tb stays on XCD tb % 8) each workgroup walks tiles tb, tb +
  // gridDim.x, ...  The one-tile form stays a separate instantiation: the
  // loop costs registers (spills) that the plain launch does not need.
  for (int tb = blockIdx.x; tb < nrows * ncolt; tb += gridDim.x) {
    int rs, ct;
    xcd_tile(tb, ncolt, nrows, &rs, &ct);
    if (rs >= nrows) continue;
    const long long m0 = (long long)rs * Tile::BM;
    const int n0 = ct * Tile::BN;
    const int mlim = (int)std::min<long long>(Tile::BM, ng - m0);
    const int nlim = std::min(Tile::BN, rk - n0);
    typename Tile::Acc acc, acc2;
    Tile::zero(acc);
    Tile::run(A1 + m0, lda, B1 + n0, ldb, K1, mlim, nlim, lds, acc);
    if constexpr (DUAL) {
      Tile::zero(acc2);
      Tile::run(A2 + m0, lda, B2 + n0, ldb, K2, mlim, nlim, lds, acc2);
    }
    const int wave = threadIdx.x >> 6;
    for (int h = 0; h < 2; h++) {
      if (wave / Tile::WGM == h) {
        if constexpr (DUAL)
          Tile::epilogue2(acc, acc2, [&](int m, int n, double v1, double v2) {
            lds[(n - h * HALF) * CS + m] = v1 * v2;
          });
        else
          Tile::epilogue(acc,
                         [&](int m, int n, double v) { lds[(n - h * HALF) * CS + m] = v * v; });
      }
      __syncthreads();
      for (int e = threadIdx.x; e < HALF * Tile::BM; e += Tile::NT) {
        const int c = e / Tile::BM, m = e % Tile::BM;
        const int n = h * HALF + c;
        if (m < mlim && n < nlim) C[(long long)(n0 + n) * ldc + m0 + m] = lds[c * CS + m];
      }
      __syncthreads();
    }
    if constexpr (!PERSIST) break;
  }
}

// Persistent form of k_zct (one product): one workgroup per CU walks the
// tiles t = blockIdx.x, blockIdx.x + gridDim.x, ...  The first K chunk of its
// next tile is fetched from HBM while the current tile's last chunk runs and
// its epilogue stores, so the CU's matrix cores wait only for the epilogue.
// gridDim.x is a multiple of 8, so tile t stays on XCD t % 8 and the column
// tiles of one row strip share that XCD's L2 as in the plain launch.
template <class Tile>
__global__ __launch_bounds__(Tile::NT, 1) void k_zct_persistent(
    int ng, int rk, const double* __restrict__ A, long long lda, int K,
    const double* __restrict__ B, int ldb, double* __restrict__ C, long long ldc) {
  __shared__ double lds[Tile::LDS_DOUBLES];
  const int nrows = (ng + Tile::BM - 1) / Tile::BM;
  const int ncolt = (rk + Tile::BN - 1) / Tile::BN;
  const int total = nrows * ncolt;
  const int nch = (K + Tile::KB - 1) / Tile::KB;
  int t = blockIdx.x;
  if (t >= total || nch == 0) return;
  auto coords = [&](int tt, long long* m0, int* n0, int* mlim, int* nlim) {
    int rs, ct;
    xcd_tile(tt, ncolt, nrows, &rs, &ct);
    *m0 = (long long)rs * Tile::BM;
    *n0 = ct * Tile::BN;
    *mlim = (int)std::min<long long>(Tile::BM, ng - *m0);
    *nlim = std::min(Tile::BN, rk - *n0);
  };
  double ra[Tile::APT], rb[Tile::BPT];
  long long m0;
  int n0, mlim, nlim;
  coords(t, &m0, &n0, &mlim, &nlim);
  Tile::load(A + m0, lda, B + n0, ldb, K, 0, mlim, nlim, ra, rb);
  Tile::store(lds, 0, ra, rb);
  __syncthreads();
  while (true) {
    const int tn = t + gridDim.x;
    const bool has_next = tn < total;
    long long m0n = 0;
    int n0n = 0, mlimn = 0, nlimn = 0;
    if (has_next) coords(tn, &m0n, &n0n, &mlimn, &nlimn);
    typename Tile::Acc acc;
    Tile::zero(acc);
    for (int c = 0; c < nch; c++) {
      const bool more = c + 1 < nch;
      if (more)
        Tile::load(A + m0, lda, B + n0, ldb, K, (c + 1) * Tile::KB, mlim, nlim, ra, rb);
      else if (has_next)
        Tile::load(A + m0n, lda, B + n0n, ldb, K, 0, mlimn, nlimn, ra, rb);
      Tile::compute(lds, c & 1, acc);
      if (more) Tile::store(lds, (c + 1) & 1, ra, rb);
      __syncthreads();
    }
    Tile::epilogue(acc, [&](int m, int n, double v) {
      if (m < mlim && n < nlim) C[(long long)(n0 + n) * ldc + m0 + m] = v * v;
    });
    if (!has_next) break;
    // Every wave is past the last chunk's barrier: buffer 0 is free.
    Tile::store(lds, 0, ra, rb);
    __syncthreads();
    t = tn;
    m0 = m0n;
    n0 = n0n;
    mlim = mlimn;
    nlim = nlimn;
  }
}

// trace summed in ascending order, eps = 1e-10 * trace / rk added to the
// diagonal (getbasis_scalapack.cpp:398-428).  One block; the diagonal passes
// through LDS in chunks of 1024 (8 KB), so the block fits on a CU beside a
// contraction workgroup that holds most of the LDS.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_trace_eps(double* __restrict__ P, int rk,
                                                   double* __restrict__ eps_out,
                                                   const CholJob* jobs) {
  constexpr int kChunk = 1024;
  __shared__ double eps_s;
  __shared__ double diag[kChunk];
  if constexpr (BATCH) {
    const CholJob& J = jobs[blockIdx.z];
    P = J.P;
    rk = J.n;
    eps_out = J.eps;
  }
  double tr = 0.0;  // thread 0's running sum
  for (int c0 = 0; c0 < rk; c0 += kChunk) {
    const int m = min(kChunk, rk - c0);
    for (int i = threadIdx.x; i < m; i += blockDim.x)
      diag[i] = P[(long long)(c0 + i) * rk + c0 + i];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = 0; i < m; i++) tr += diag[i];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
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

// Value of lane `src` (a compile-time constant after unrolling) in every lane,
// through two v_readlane_b32: no LDS, no barrier.
__device__ __forceinline__ double wave_bcast(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

constexpr int kDiagCW = kNB / 4;  // columns (and inverse rows) per wave of the diagonal factor

// Interval J of the diagonal-block factor and inverse, run in lockstep.
// Wave w keeps columns [16w, 16w+16) of every row (lane r = row r) of the
// factor in v, and rows [16w, 16w+16) of column `lane` of the inverse in x.
// In interval J (J = 0..64, one barrier each):
//  * every wave applies factor step J-1, v(:,k) -= L(r,J-1) L(k,J-1), to all
//    its columns with no mask: L(g,J-1) = 0 for g <= J-1 (those lanes publish
//    0), so finished columns are unchanged;
//  * every wave applies inverse step J-2, x_i -= L(i,J-2) x_{J-2}, likewise
//    (L(i,J-2) = 0 for i <= J-2);
//  * the owner of column J forms it (pivot, sqrt, correctly rounded
//    divisions) and publishes L(:,J) and L_JJ; a non-positive pivot is
//    recorded in *fail_s (first failing step + 1) and replaced by 1 so that
//    the rest stays finite (nothing is written back then);
//  * the owner of inverse row J-1 divides it by L_{J-1,J-1} and publishes it.
// L(:,J) lives in col[J mod 3] from interval J to J+2, x_T in xt[T mod 2]
// from interval T+1 to T+2.
template <int J>
__device__ __forceinline__ void potrf_diag_interval(double (&v)[kDiagCW], double (&x)[kDiagCW],
                                                    int lane, int w, int* fail_s, double* col,
                                                    double* ldiag, double* xt) {
  if constexpr (J >= 1 && J <= kNB) {
    const double* c = col + ((J - 1) % 3) * kNB;
    const double a = c[lane];
    const double* cw = c + w * kDiagCW;
#pragma unroll
    for (int k = 0; k < kDiagCW; k += 2) {
      const double2 l = *reinterpret_cast<const double2*>(cw + k);
      v[k] -= a * l.x;
      v[k + 1] -= a * l.y;
    }
  }
  if constexpr (J >= 2 && J <= kNB + 1) {
    constexpr int T = J - 2;
    const double xT = xt[(T & 1) * kNB + lane];
    const double* lw = col + (T % 3) * kNB + w * kDiagCW;
#pragma unroll
    for (int i = 0; i < kDiagCW; i += 2) {
      const double2 q = *reinterpret_cast<const double2*>(lw + i);
      x[i] -= q.x * xT;
      x[i + 1] -= q.y * xT;
    }
  }
  if constexpr (J < kNB) {
    constexpr int o = J / kDiagCW, jj = J % kDiagCW;
    if (w == o) {
      double piv = wave_bcast(v[jj], J);
      const bool bad = !(piv > 0.0);
      if (bad && *fail_s == 0) *fail_s = J + 1;
      piv = bad ? 1.0 : piv;
      const double d = sqrt(piv);
      const double q = v[jj] / d;
      const double lij = lane > J ? q : 0.0;
      v[jj] = lane == J ? d : lij;
      ldiag[J] = d;
      col[(J % 3) * kNB + lane] = lij;
    }
  }
  if constexpr (J >= 1 && J <= kNB) {
    constexpr int T = J - 1, o = T / kDiagCW, tt = T % kDiagCW;
    if (w == o) {
      x[tt] = x[tt] / ldiag[T];
      xt[(T & 1) * kNB + lane] = x[tt];
    }
  }
  __syncthreads();
}

template <int... J>
__device__ __forceinline__ void potrf_diag_intervals(double (&v)[kDiagCW], double (&x)[kDiagCW],
                                                     int lane, int w, int* fail_s, double* col,
                                                     double* ldiag, double* xt,
                                                     std::integer_sequence<int, J...>) {
  (potrf_diag_interval<J>(v, x, lane, w, fail_s, col, ldiag, xt), ...);
}

// Factor the jb x jb diagonal block at J0, write L_JJ back and its inverse
// into Dinv[J] (64 x 64, column-major, zero padded).  Four waves, one per
// SIMD: lane r holds row r, wave w its columns [16w, 16w+16) (static indices:
// every interval is its own instance), so one step's products are spread over
// the four SIMDs.  Each entry takes its updates in ascending j and each
// division is correctly rounded, so the factor is the one the unblocked
// algorithm gives; the entries right of the diagonal (k > r) are updated too
// (never read), and the finished columns take exact zero updates (the sign
// of an exact zero aside, they are unchanged).  The inverse is formed
// right-looking, column `lane` of inv(L_JJ), rows split over the waves, one
// step behind the factor: each entry accumulates its terms in ascending t
// before the division by L_tt.  Past jb the block is the identity.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_potrf_diag(double* __restrict__ A, int n, int J0,
                                                    double* __restrict__ Dinv,
                                                    int* __restrict__ info, const CholJob* jobs) {
  __shared__ __align__(16) double col[3 * kNB];
  __shared__ __align__(16) double xt[2 * kNB];
  __shared__ double ldiag[kNB];
  __shared__ int fail_s;
  if constexpr (BATCH) {
    const CholJob& J = jobs[blockIdx.z];
    A = J.P;
    n = J.n;
    Dinv = J.Dinv;
    info = J.info;
    if (J0 >= n) return;
  }
  if (*info != 0) return;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int jb = min(kNB, n - J0);
  const double* Ab = A + (long long)J0 * n + J0;
  if (threadIdx.x == 0) fail_s = 0;
  double v[kDiagCW];  // v[k] = A[lane][16w + k] on and below the diagonal
  double x[kDiagCW];  // x[k] = inv(L_JJ)[16w + k][lane], from e_lane
#pragma unroll
  for (int k = 0; k < kDiagCW; k++) {
    const int g = w * kDiagCW + k;
    v[k] = (g < jb && lane < jb && lane >= g) ? Ab[(long long)g * n + lane] : (g == lane ? 1.0 : 0.0);
    x[k] = g == lane ? 1.0 : 0.0;
  }
  __syncthreads();
  potrf_diag_intervals(v, x, lane, w, &fail_s, col, ldiag, xt,
                       std::make_integer_sequence<int, kNB + 1>{});
  const int fail = fail_s;
  if (fail) {
    if (threadIdx.x == 0) atomicCAS(info, 0, J0 + fail);
    return;
  }
  double* Aw = A + (long long)J0 * n + J0;
  double* D = Dinv + (long long)(J0 / kNB) * kNB * kNB;
#pragma unroll
  for (int k = 0; k < kDiagCW; k++) {
    const int g = w * kDiagCW + k;
    if (g < jb && lane < jb && lane >= g) Aw[(long long)g * n + lane] = v[k];
    D[(long long)lane * kNB + g] = (lane < jb && g < jb) ? x[k] : 0.0;
  }
}

// The trailing update of one step is a rank-64 product, so the work per tile
// is small and the step's latency is set by how many tiles run side by side:
// 32 x 32 tiles (128 x 128 left most CUs idle, and 64 x 64 fill at most 120
// CUs at rk 1024).  Every entry sums the same k range in the same 16-wide
// chunks whatever the tile.  The panel's 1-D grid over rows needs tiles at
// least kNB wide.
typedef MfmaTile<64, 64, 16, 2, 2> PanelTile;
typedef MfmaTile<32, 32, 16, 2, 2> SyrkTile;
static_assert(PanelTile::BN >= kNB, "the panel tile must cover the block's columns");

// L_IJ = A_IJ inv(L_JJ)^T for the rows below the diagonal block (MFMA).
// Each workgroup owns PanelTile::BM rows, so the in-place overwrite is safe once its
// GEMM has read them.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_potrf_panel(double* __restrict__ A, int n, int J0,
                                                     const double* __restrict__ Dinv,
                                                     const int* __restrict__ info,
                                                     const CholJob* jobs) {
  __shared__ double lds[PanelTile::LDS_DOUBLES];
  if constexpr (BATCH) {
    const CholJob& J = jobs[blockIdx.z];
    A = J.P;
    n = J.n;
    Dinv = J.Dinv;
    info = J.info;
    if (J0 >= n) return;
  }
  if (*info != 0) return;
  const int jb = min(kNB, n - J0);
  const int I0 = J0 + jb + blockIdx.x * PanelTile::BM;
  if (I0 >= n) return;
  const int mlim = min(PanelTile::BM, n - I0);
  double* Ai = A + (long long)J0 * n + I0;
  const double* D = Dinv + (long long)(J0 / kNB) * kNB * kNB;
  typename PanelTile::Acc acc;
  PanelTile::zero(acc);
  // B(k, c) = inv(L_JJ)(c, k) = D[k*64 + c]: k-major with ld 64.
  PanelTile::run(Ai, n, D, kNB, jb, mlim, jb, lds, acc);
#pragma unroll
  for (int j = 0; j < PanelTile::TN; j++) {
    const int c = PanelTile::col(j);
    if (c >= jb) continue;
#pragma unroll
    for (int i = 0; i < PanelTile::TM; i++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = PanelTile::row(i, r);
        if (m < mlim) Ai[(long long)c * n + m] = acc[i][j][r];
      }
  }
}

// Trailing update A_IK -= L_IJ L_KJ^T on the lower tiles (MFMA).
template <bool BATCH>
__global__ __launch_bounds__(256) void k_potrf_syrk(double* __restrict__ A, int n, int J0,
                                                    const int* __restrict__ info,
                                                    const CholJob* jobs) {
  __shared__ double lds[SyrkTile::LDS_DOUBLES];
  if constexpr (BATCH) {
    const CholJob& J = jobs[blockIdx.z];
    A = J.P;
    n = J.n;
    info = J.info;
    if (J0 >= n) return;
  }
  if (*info != 0) return;
  const int ti = blockIdx.x, tk = blockIdx.y;
  if (tk > ti) return;
  const int jb = min(kNB, n - J0);
  const int T0 = J0 + jb;
  const int I0 = T0 + ti * SyrkTile::BM, K0 = T0 + tk * SyrkTile::BN;
  if (I0 >= n) return;
  const int mlim = min(SyrkTile::BM, n - I0), nlim = min(SyrkTile::BN, n - K0);
  typename SyrkTile::Acc acc;
  SyrkTile::zero(acc);
  // A(m, k) = L[I0+m, J0+k]; B(k, c) = L[K0+c, J0+k] (k-major, ld n).
  SyrkTile::run(A + (long long)J0 * n + I0, n, A + (long long)J0 * n + K0, n, jb, mlim, nlim, lds,
                acc);
#pragma unroll
  for (int j = 0; j < SyrkTile::TN; j++) {
    const int c = SyrkTile::col(j);
    if (c >= nlim) continue;
    double* Ac = A + (long long)(K0 + c) * n + I0;
#pragma unroll
    for (int i = 0; i < SyrkTile::TM; i++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = SyrkTile::row(i, r);
        if (m < mlim && I0 + m >= K0 + c) Ac[m] -= acc[i][j][r];
      }
  }
}

// ---------------------------------------------------------------------------
// Theta = Z (L L^T)^{-1} in place, in block columns of kSB = 128.
//
// Forward (X L^T = B):  X_J = [X_0 .. X_{J-1}, B_J] F_J^T with
//   F_JQ = -inv(L_JJ) L_JQ (Q < J), F_JJ = inv(L_JJ);  Ft holds F column-major.
// Backward (Theta L = X): Theta_J = [X_J, Theta_{J+1} ..] G_J with
//   G_PJ = -L_PJ inv(L_JJ) (P > J), G_JJ = inv(L_JJ);  Gt holds G row-major.
// Each block column is then one GEMM over already-solved columns; a workgroup
// owns a strip of rows and walks the block columns in order.  Block inverses of
// 128 x 128 come from the Cholesky's 64 x 64 ones by the 2 x 2 block formula.
// ---------------------------------------------------------------------------

// Dinv128[J] = inv(L_JJ) for the 128-blocks: [[Da, 0], [-Db L_ba Da, Db]].
// Both 64 x 64 products are register-blocked, 4 x 4 outputs per thread, with
// k ascending over the full range: the terms outside the triangles are exact
// zeros (Da and Db are zero above the diagonal), added to a sum that is still
// zero (T) or added last (X), so each entry is the same sequence of roundings
// as the triangular sums T(r,c) = sum_{k>=c} and X(r,c) = -sum_{k<=r}.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_dinv128(const double* __restrict__ L, int n,
                                                 const double* __restrict__ Dinv64,
                                                 double* __restrict__ Dinv128,
                                                 const CholJob* jobs) {
  __shared__ __align__(16) double daT[kNB][kNB];  // daT[k][c] = Da(k, c)
  __shared__ __align__(16) double dbT[kNB][kNB];  // dbT[k][r] = Db(r, k)
  __shared__ __align__(16) double lt[kNB][kNB];   // lt[k][r] = L_ba(r, k); then T row-major
  const int J = blockIdx.x, J0 = J * kSB;
  if constexpr (BATCH) {
    const CholJob& jb = jobs[blockIdx.z];
    L = jb.P;
    n = jb.n;
    Dinv64 = jb.Dinv;
    Dinv128 = jb.Dinv128;
    if (J0 >= n) return;
  }
  const int t = threadIdx.x;
  const double* Da = Dinv64 + (long long)(2 * J) * kNB * kNB;
  const double* Db = Dinv64 + (long long)(2 * J + 1) * kNB * kNB;
  const bool hasb = J0 + kNB < n;
  double* D = Dinv128 + (long long)J * kSB * kSB;
  for (int e = t; e < kNB * kNB; e += 256) {
    const int r = e % kNB, c = e / kNB;  // Da, Db column-major: e = c*64 + r
    daT[r][c] = Da[e];
    dbT[c][r] = hasb ? Db[e] : 0.0;
    // L_ba(r, c) = L[J0+64+r, J0+c]
    lt[c][r] = (hasb && J0 + kNB + r < n) ? L[(long long)(J0 + c) * n + J0 + kNB + r] : 0.0;
  }
  __syncthreads();
  const int r0 = 4 * (t % 16), c0 = 4 * (t / 16);
  double acc[4][4];
  // T = L_ba Da: T(r, c) = sum_k L_ba(r, k) Da(k, c).
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
#pragma unroll 4
  for (int k = 0; k < kNB; k++) {
    const double2 a01 = *reinterpret_cast<const double2*>(&lt[k][r0]);
    const double2 a23 = *reinterpret_cast<const double2*>(&lt[k][r0 + 2]);
    const double2 b01 = *reinterpret_cast<const double2*>(&daT[k][c0]);
    const double2 b23 = *reinterpret_cast<const double2*>(&daT[k][c0 + 2]);
    const double a[4] = {a01.x, a01.y, a23.x, a23.y}, b[4] = {b01.x, b01.y, b23.x, b23.y};
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) acc[i][j] += a[i] * b[j];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) lt[r0 + i][c0 + j] = acc[i][j];  // T row-major
  __syncthreads();
  // X = -Db T: X(r, c) = -sum_k Db(r, k) T(k, c).
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
#pragma unroll 4
  for (int k = 0; k < kNB; k++) {
    const double2 a01 = *reinterpret_cast<const double2*>(&dbT[k][r0]);
    const double2 a23 = *reinterpret_cast<const double2*>(&dbT[k][r0 + 2]);
    const double2 b01 = *reinterpret_cast<const double2*>(&lt[k][c0]);
    const double2 b23 = *reinterpret_cast<const double2*>(&lt[k][c0 + 2]);
    const double a[4] = {a01.x, a01.y, a23.x, a23.y}, b[4] = {b01.x, b01.y, b23.x, b23.y};
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) acc[i][j] += a[i] * b[j];
  }
  for (int e = t; e < kSB * kSB; e += 256) {
    const int rr = e % kSB, cc = e / kSB;
    double v = 0.0;
    if (rr < kNB && cc < kNB) v = daT[rr][cc];
    else if (rr >= kNB && cc >= kNB) v = dbT[cc - kNB][rr - kNB];
    D[e] = v;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int i = 0; i < 4; i++) D[(long long)(c0 + j) * kSB + kNB + r0 + i] = -acc[i][j];
}

typedef MfmaTile<128, 128, 16, 2, 2, false> FgTile;

// One workgroup per block pair (P, Q), Q <= P: F_PQ and G_PQ.
template <bool BATCH>
__global__ __launch_bounds__(256, 1) void k_make_fg(const double* __restrict__ L, int n,
                                                    const double* __restrict__ Dinv128,
                                                    double* __restrict__ Ft,
                                                    double* __restrict__ Gt,
                                                    const CholJob* jobs) {
  __shared__ double lds[FgTile::LDS_DOUBLES];
  const int P = blockIdx.x, Q = blockIdx.y;
  if (Q > P) return;
  if constexpr (BATCH) {
    const CholJob& J = jobs[blockIdx.z];
    L = J.P;
    n = J.n;
    Dinv128 = J.Dinv128;
    Ft = J.Ft;
    Gt = J.Gt;
    if (P * kSB >= n) return;
  }
  const int P0 = P * kSB, Q0 = Q * kSB;
  const int pb = min(kSB, n - P0), qb = min(kSB, n - Q0);
  const double* DP = Dinv128 + (long long)P * kSB * kSB;
  const double* DQ = Dinv128 + (long long)Q * kSB * kSB;
  if (P == Q) {
    for (int e = threadIdx.x; e < kSB * kSB; e += 256) {
      const int i = e % kSB, j = e / kSB;
      if (i < pb && j < qb) {
        const double v = DP[e];
        Ft[(long long)(Q0 + j) * n + P0 + i] = v;
        Gt[(long long)(P0 + i) * n + Q0 + j] = v;
      }
    }
    return;
  }
  typename FgTile::Acc acc;
  // F_PQ = -Dinv_P L_PQ: A = Dinv_P (ld 128), B(k, j) = L[P0+k, Q0+j] (k contiguous, ld n).
  FgTile::zero(acc);
  FgTile::run(DP, kSB, L + (long long)Q0 * n + P0, n, pb, pb, qb, lds, acc);
#pragma unroll
  for (int j = 0; j < FgTile::TN; j++) {
    const int jj = FgTile::col(j);
#pragma unroll
    for (int i = 0; i < FgTile::TM; i++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int ii = FgTile::row(i, r);
        if (ii < pb && jj < qb) Ft[(long long)(Q0 + jj) * n + P0 + ii] = -acc[i][j][r];
      }
  }
  // G_PQ = -L_PQ Dinv_Q: A(i, k) = L[P0+i, Q0+k] (ld n), B(k, j) = Dinv_Q[k][j] (k contiguous, ld 128).
  FgTile::zero(acc);
  FgTile::run(L + (long long)Q0 * n + P0, n, DQ, kSB, qb, pb, qb, lds, acc);
#pragma unroll
  for (int j = 0; j < FgTile::TN; j++) {
    const int jj = FgTile::col(j);
#pragma unroll
    for (int i = 0; i < FgTile::TM; i++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int ii = FgTile::row(i, r);
        if (ii < pb && jj < qb) Gt[(long long)(P0 + ii) * n + Q0 + jj] = -acc[i][j][r];
      }
  }
}

// The in-place solve.  A workgroup owns Tile::BM rows; Tile::BN == kSB.
// unit_lower: T is the identity (ng == n) and only the forward pass runs, so
// T becomes L^{-T} (upper triangular).  Row i of it (column i of L^{-1})
// vanishes left of column i: block columns before the strip's own are zero
// and stay so, and the forward GEMMs start there.
// LDS_EPI: each block column is written back through LDS in two halves of
// 64 columns, as contiguous column segments (needs a tile with WGN == 2).
template <class Tile, bool BATCH, bool LDS_EPI = false>
__global__ __launch_bounds__(Tile::NT, 1) void k_potrs(double* __restrict__ T, long long ng,
                                                       long long ldt, int n,
                                                       const double* __restrict__ Ft,
                                                       const double* __restrict__ Gt,
                                                       int unit_lower, const CholJob* jobs) {
  static_assert(Tile::BN == kSB, "block column width");
  static_assert(kSB % Tile::BM == 0 || Tile::BM % kSB == 0, "strip / block alignment");
  constexpr int HALF = Tile::BN / 2, CS = Tile::BM + 4, STAGE = LDS_EPI ? HALF * CS : 0;
  static_assert(!LDS_EPI || Tile::WGN == 2, "one wave column per half");
  __shared__ double lds[Tile::LDS_DOUBLES > STAGE ? Tile::LDS_DOUBLES : STAGE];
  const long long m0 = (long long)blockIdx.x * Tile::BM;
  if constexpr (BATCH) {  // T = Gt of the instance, which holds the identity
    const CholJob& J = jobs[blockIdx.z];
    T = J.Gt;
    n = J.n;
    ng = ldt = J.n;
    Ft = J.Ft;
    Gt = J.Gt;
    if (m0 >= ng) return;
  }
  const int mlim = (int)std::min<long long>(Tile::BM, ng - m0);
  double* Tm = T + m0;
  const int nJ = (n + kSB - 1) / kSB;
  const int Js = unit_lower ? (int)(m0 / kSB) : 0;
  const int npass = unit_lower ? 1 : 2;
  for (int pass = 0; pass < npass; pass++) {
    for (int s = pass == 0 ? Js : 0; s < nJ; s++) {
      const int J = pass == 0 ? s : nJ - 1 - s;
      const int J0 = J * kSB;
      const int jb = min(kSB, n - J0);
      const int kbeg = pass == 0 ? Js * kSB : J0;
      const int kend = pass == 0 ? J0 + jb : n;
      typename Tile::Acc acc;
      Tile::zero(acc);
      const double* Bop = pass == 0 ? Ft : Gt;
      Tile::run(Tm + (long long)kbeg * ldt, ldt, Bop + (long long)kbeg * n + J0, n, kend - kbeg,
                mlim, jb, lds, acc);
      // Tile::run ends with a barrier: every read of block column J is done.
      if constexpr (LDS_EPI) {
        const int wave = threadIdx.x >> 6;
        for (int h = 0; h < 2; h++) {
          if (wave / Tile::WGM == h)
            Tile::epilogue(acc, [&](int m, int c, double v) { lds[(c - h * HALF) * CS + m] = v; });
          __syncthreads();
          for (int e = threadIdx.x; e < HALF * Tile::BM; e += Tile::NT) {
            const int c = h * HALF + e / Tile::BM, m = e % Tile::BM;
            if (m < mlim && c < jb) Tm[(long long)(J0 + c) * ldt + m] = lds[(e / Tile::BM) * CS + m];
          }
          __syncthreads();
        }
      } else {
        Tile::epilogue(acc, [&](int m, int c, double v) {
          if (m < mlim && c < jb) Tm[(long long)(J0 + c) * ldt + m] = v;
        });
        // The next block column reads what this one wrote.
        __syncthreads();
      }
    }
  }
}

// 32 x 32 tiles: the product sits on the mixed forms' critical path, and its
// time is that of the tiles of the first block row (k runs over all of rk):
// at rk 1024 the lower triangle is 528 tiles (136 of 64 x 64, 36 of 128 x 128).
typedef MfmaTile<32, 32, 16, 2, 2> XxtTile;

// W = X X^T for the upper-triangular X = L^{-T} (n x n, column-major), so
// W = L^{-T} L^{-1} = (L L^T)^{-1}.  One workgroup per lower tile (ti >= tj);
// X(i, k) = 0 for k < i, so the sum starts at the tile's first row; the tile
// is written at (i, j) and mirrored to (j, i).
template <bool BATCH>
__global__ __launch_bounds__(XxtTile::NT, 1) void k_xxt(const double* __restrict__ X, int n,
                                                        double* __restrict__ W,
                                                        const CholJob* jobs) {
  __shared__ double lds[XxtTile::LDS_DOUBLES];
  if constexpr (BATCH) {
    const CholJob& J = jobs[blockIdx.z];
    X = J.Gt;
    W = J.W;
    n = J.n;
  }
  const int ti = blockIdx.x, tj = blockIdx.y;
  if (tj > ti) return;
  const int i0 = ti * XxtTile::BM, j0 = tj * XxtTile::BN;
  if (i0 >= n) return;
  const int mlim = min(XxtTile::BM, n - i0), nlim = min(XxtTile::BN, n - j0);
  typename XxtTile::Acc acc;
  XxtTile::zero(acc);
  XxtTile::run(X + (long long)i0 * n + i0, n, X + (long long)i0 * n + j0, n, n - i0, mlim, nlim,
               lds, acc);
  XxtTile::epilogue(acc, [&](int m, int c, double v) {
    if (m < mlim && c < nlim) {
      W[(long long)(j0 + c) * n + i0 + m] = v;
      W[(long long)(i0 + m) * n + j0 + c] = v;
    }
  });
}

// The in-place solve of k_potrs on the FP64 vector FMA pipe (DppGemm: B
// broadcast by DPP, operands staged by LDS DMA).  A workgroup owns BM = 64 TM
// rows and walks the block columns as k_potrs does.  Requires ng a multiple
// of BM (every strip whole), n a multiple of KB, 16-byte aligned rows (ldt
// even, T, Ft, Gt 16-byte aligned); columns of a B row past the block width
// are read (the F / G allocations carry one block width of padding) and feed
// only output columns that are not stored.
template <int TM, int TG, int KB, int WPE>
__global__ __launch_bounds__((DppGemm<TM, TG, KB>::NT), WPE) void k_potrs_dpp(
    double* __restrict__ T, long long ldt, int n, const double* __restrict__ Ft,
    const double* __restrict__ Gt) {
  typedef DppGemm<TM, TG, KB> G;
  static_assert(G::BN == kSB, "block column width");
  __shared__ double lds[G::LDS_DOUBLES];
  double* Tm = T + (long long)blockIdx.x * G::BM;
  const int nJ = (n + kSB - 1) / kSB;
  for (int pass = 0; pass < 2; pass++) {
    for (int s = 0; s < nJ; s++) {
      const int J = pass == 0 ? s : nJ - 1 - s;
      const int J0 = J * kSB;
      const int jb = min(kSB, n - J0);
      const int kbeg = pass == 0 ? 0 : J0;
      const int K = (pass == 0 ? J0 + jb : n) - kbeg;
      typename G::Acc acc;
      G::zero(acc);
      G::run(Tm + (long long)kbeg * ldt, ldt, (pass == 0 ? Ft : Gt) + (long long)kbeg * n + J0, n,
             K, lds, acc);
      // Every global read of this block column (block J's own columns among
      // them) was a copy that had landed by the last chunk's barrier, so the
      // columns can be overwritten while other waves still run FMAs from LDS.
#pragma unroll
      for (int g = 0; g < TG; g++)
#pragma unroll
        for (int j = 0; j < 16; j++) {
          const int col = G::col(g, j);
          if (col < jb) {
#pragma unroll
            for (int i = 0; i < TM; i++)
              Tm[(long long)(J0 + col) * ldt + G::row(i)] = acc[(i * TG + g) * 16 + j];
          }
        }
      // The next block column reads what this one wrote.
      __syncthreads();
    }
  }
}

// Z C^T of k_zct on the FP64 vector FMA pipe (DppGemm): C = P1 o P1 or P1 o P2
// with P1 = A1 B1 (K1 deep), P2 = A2 B2 (K2 deep), one BM x 128 tile per
// workgroup step (a grid smaller than the tile count walks tiles tb,
// tb + gridDim.x, ..., a multiple of 8 apart so that a tile stays on its XCD).
// The factors live only in accumulators; the epilogue writes their product.
// Requires ng a multiple of BM, rk a multiple of 128, K1 and K2 multiples of
// KB, even lda and ldb, 16-byte aligned operands.
template <int TM, int TG, int KB, int WPE, bool DUAL>
__global__ __launch_bounds__((DppGemm<TM, TG, KB>::NT), WPE) void k_zct_dpp(
    int ng, int rk, const double* __restrict__ A1, const double* __restrict__ A2,
    long long lda, int K1, int K2, const double* __restrict__ B1,
    const double* __restrict__ B2, int ldb, double* __restrict__ C, long long ldc) {
  typedef DppGemm<TM, TG, KB> G;
  __shared__ double lds[G::LDS_DOUBLES];
  const int nrows = ng / G::BM;
  const int ncolt = rk / G::BN;
  for (int tb = blockIdx.x; tb < nrows * ncolt; tb += gridDim.x) {
    int rs, ct;
    xcd_tile(tb, ncolt, nrows, &rs, &ct);
    const long long m0 = (long long)rs * G::BM;
    const int n0 = ct * G::BN;
    typename G::Acc acc, acc2;
    G::zero(acc);
    G::run(A1 + m0, lda, B1 + n0, ldb, K1, lds, acc);
    if constexpr (DUAL) {
      G::zero(acc2);
      G::run(A2 + m0, lda, B2 + n0, ldb, K2, lds, acc2);
    }
#pragma unroll
    for (int g = 0; g < TG; g++)
#pragma unroll
      for (int j = 0; j < 16; j++)
#pragma unroll
        for (int i = 0; i < TM; i++) {
          const int e = (i * TG + g) * 16 + j;
          const double v = acc[e];
          C[(long long)(n0 + G::col(g, j)) * ldc + m0 + G::row(i)] = DUAL ? v * acc2[e] : v * v;
        }
  }
}

// Tile variants of the solve (chosen per shape, see launch_potrs).
typedef MfmaTile<256, kSB, 16, 4, 2> PotrsTile0;  // 8 waves of 64 x 64
typedef MfmaTile<256, kSB, 16, 4, 1> PotrsTile1;  // 4 waves of 64 x 128
typedef MfmaTile<128, kSB, 16, 2, 2> PotrsTile2;  // 4 waves of 64 x 64
typedef ValuTile<256, kSB, 16, 1, 8> PotrsTile3;  // vector FMA, 8 waves of 256 x 16
typedef MfmaTile<32, kSB, 16, 2, 2> PotrsTile4;   // 4 waves of 16 x 64: few rows, many workgroups
typedef MfmaTile<16, kSB, 16, 1, 4> PotrsTile5;   // 4 waves of 16 x 32
typedef ScalarBTile<256, kSB, 8> PotrsTile7;  // vector FMA, B in scalar registers, no LDS
typedef ScalarBTile<128, kSB, 8> PotrsTile8;  // the same on 128-row strips

}  // namespace

int launch_gather_psimu(const double* psi, long long ld, int nstate, const int* piv, int rk,
                        double* psimuT, hipStream_t st) {
  hipLaunchKernelGGL(k_gather, dim3(cdiv((long long)rk * nstate, 256)), dim3(256), 0, st, psi, ld,
                     nstate, piv, rk, psimuT);
  ISDF_HIP_LAUNCHED("k_gather");
  return 0;
}

int zct_variants() { return 7; }

int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cus[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      v = 256;
    cus[dev] = v;
  }
  return cus[dev];
}

// The MFMA tile for a grid of ng points: 4 (LDS-staged epilogue) measured 6-9% faster
// than 0 on the 64^3 and 96^3 workloads.
int zct_mfma_variant(long long ng) { return ng >= 256LL * 256 ? 4 : 2; }

int zct_pick_variant(long long ng, int requested) {
  // Default: the vector-FMA product (5), which launch_zct replaces by the MFMA tile
  // where its shape conditions fail.  Per launch at 96^3 18.0 against 19.4 ms, on the
  // dual 96^3 workload 66 against 74 ms, at 64^3 1.6 against 2.1 ms, bitwise equal.
  return requested >= 0 ? requested : 5;
}

int grid_leaving_cus(long long tiles, int reserve) {
  if (reserve <= 0) return (int)tiles;
  return (int)std::min<long long>(tiles, std::max(8, (device_cus() - reserve) / 8 * 8));
}

namespace {
// Whether k_zct_dpp with BM-row strips and KB-deep chunks can take the product.
bool dpp_zct_fits(int ng, int rk, long long lda, int ldb, int K1, int K2, bool dual,
                  const double* A1, const double* A2, const double* B1, const double* B2, int bm,
                  int kb) {
  const uintptr_t al = (uintptr_t)A1 | (uintptr_t)B1 | (dual ? (uintptr_t)A2 | (uintptr_t)B2 : 0);
  return ng % bm == 0 && rk % 128 == 0 && K1 % kb == 0 && (!dual || K2 % kb == 0) &&
         lda % 2 == 0 && ldb % 2 == 0 && al % 16 == 0;
}

// k_zct_dpp shapes: (TM, TG, KB, waves per SIMD); WGS workgroups per CU fit
// at that occupancy.  With CUs reserved for the concurrent Cholesky branch
// the grid is WGS per remaining CU (a multiple of 8) and walks the tiles.
template <int TM, int TG, int KB, int WPE, int WGS, bool DUAL>
int launch_zct_dpp(int ng, int rk, const double* A1, const double* A2, long long lda, int K1,
                   int K2, const double* B1, const double* B2, int ldb, double* C, long long ldc,
                   hipStream_t st, int reserve_cus) {
  typedef DppGemm<TM, TG, KB> G;
  if (!dpp_zct_fits(ng, rk, lda, ldb, K1, K2, DUAL, A1, A2, B1, B2, G::BM, KB))
    return set_error(1, "the vector-FMA Z C^T needs whole row strips, rk a multiple of 128, K "
                        "a multiple of its k chunk and 16-byte aligned operands");
  const long long tiles = (long long)(ng / G::BM) * (rk / G::BN);
  long long grid = tiles;
  if (reserve_cus > 0)
    grid = std::min<long long>(tiles, (long long)WGS * std::max(8, (device_cus() - reserve_cus) / 8 * 8));
  hipLaunchKernelGGL((k_zct_dpp<TM, TG, KB, WPE, DUAL>), dim3((unsigned)grid), dim3(G::NT), 0, st,
                     ng, rk, A1, A2, lda, K1, K2, B1, B2, ldb, C, ldc);
  return 0;
}
}  // namespace

int launch_zct(const double* psi, long long ld, int ng, const double* psimuT, int rk, int s1,
               int K1, int s2, int K2, bool dual, double* C, long long ldc, int variant,
               hipStream_t st, int reserve_cus) {
  const double* A1 = psi + (long long)s1 * ld;
  const double* A2 = psi + (long long)s2 * ld;
  const double* B1 = psimuT + (long long)s1 * rk;
  const double* B2 = psimuT + (long long)s2 * rk;
  variant = zct_pick_variant(ng, variant);
  if ((variant == 5 || variant == 6) &&
      !dpp_zct_fits(ng, rk, ld, rk, K1, K2, dual, A1, A2, B1, B2, variant == 5 && dual ? 64 : 128,
                    variant == 6 && !dual ? 16 : 8))
    variant = zct_mfma_variant(ng);
  if (variant == 5 || variant == 6) {
    int rc;
    if (!dual)
      rc = variant == 5 ? launch_zct_dpp<2, 2, 8, 3, 3, false>(ng, rk, A1, A2, ld, K1, K2, B1, B2,
                                                               rk, C, ldc, st, reserve_cus)
                        : launch_zct_dpp<2, 1, 16, 4, 2, false>(ng, rk, A1, A2, ld, K1, K2, B1,
                                                                B2, rk, C, ldc, st, reserve_cus);
    else
      rc = variant == 5 ? launch_zct_dpp<1, 2, 8, 3, 3, true>(ng, rk, A1, A2, ld, K1, K2, B1, B2,
                                                              rk, C, ldc, st, reserve_cus)
                        : launch_zct_dpp<2, 1, 8, 2, 1, true>(ng, rk, A1, A2, ld, K1, K2, B1, B2,
                                                              rk, C, ldc, st, reserve_cus);
    if (rc) return rc;
    ISDF_HIP_LAUNCHED("k_zct_dpp");
    return 0;
  }
#define ISDF_ZCT_LAUNCH(TL, DUALV)                                                              \
  hipLaunchKernelGGL((k_zct<TL, DUALV>), dim3(cdiv(ng, TL::BM) * cdiv(rk, TL::BN)), dim3(TL::NT), \
                     0, st, ng, rk, A1, A2, ld, K1, K2, B1, B2, rk, C, ldc)
  if (!dual) {
    if (variant == 0) ISDF_ZCT_LAUNCH(ZctTile0, false);
    else if (variant == 1) ISDF_ZCT_LAUNCH(ZctTile1, false);
    else if (variant == 4) {
      const long long tiles = (long long)cdiv(ng, ZctTile0::BM) * cdiv(rk, ZctTile0::BN);
      const int grid = grid_leaving_cus(tiles, reserve_cus);
      if (grid < tiles)
        hipLaunchKernelGGL((k_zct_lds_epilogue<ZctTile0, false, true>), dim3(grid),
                           dim3(ZctTile0::NT), 0, st, ng, rk, A1, A2, ld, K1, K2, B1, B2, rk, C,
                           ldc);
      else
        hipLaunchKernelGGL((k_zct_lds_epilogue<ZctTile0, false>), dim3(grid), dim3(ZctTile0::NT),
                           0, st, ng, rk, A1, A2, ld, K1, K2, B1, B2, rk, C, ldc);
    } else if (variant == 3) {
      const long long total = (long long)cdiv(ng, ZctTile0::BM) * cdiv(rk, ZctTile0::BN);
      const int grid = (int)std::min<long long>(total, std::max(8, device_cus() / 8 * 8));
      hipLaunchKernelGGL(k_zct_persistent<ZctTile0>, dim3(grid), dim3(ZctTile0::NT), 0, st, ng,
                         rk, A1, ld, K1, B1, rk, C, ldc);
    } else ISDF_ZCT_LAUNCH(ZctTile2, false);
  } else {
    if (variant == 2) ISDF_ZCT_LAUNCH(ZctTileDual1, true);
    else if (variant == 4) {
      const long long tiles = (long long)cdiv(ng, ZctTileDual0::BM) * cdiv(rk, ZctTileDual0::BN);
      const int grid = grid_leaving_cus(tiles, reserve_cus);
      if (grid < tiles)
        hipLaunchKernelGGL((k_zct_lds_epilogue<ZctTileDual0, true, true>), dim3(grid),
                           dim3(ZctTileDual0::NT), 0, st, ng, rk, A1, A2, ld, K1, K2, B1, B2, rk,
                           C, ldc);
      else
        hipLaunchKernelGGL((k_zct_lds_epilogue<ZctTileDual0, true>), dim3(grid),
                           dim3(ZctTileDual0::NT), 0, st, ng, rk, A1, A2, ld, K1, K2, B1, B2, rk,
                           C, ldc);
    }
    else ISDF_ZCT_LAUNCH(ZctTileDual0, true);
  }
#undef ISDF_ZCT_LAUNCH
  ISDF_HIP_LAUNCHED("k_zct");
  return 0;
}

int launch_trace_eps(double* P, int rk, double* eps_dev, hipStream_t st) {
  hipLaunchKernelGGL(k_trace_eps<false>, dim3(1), dim3(256), 0, st, P, rk, eps_dev,
                     (const CholJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_trace_eps");
  return 0;
}

int launch_potrf(double* P, int rk, double* Dinv, int* info_dev, hipStream_t st) {
  const int nJ = (rk + kNB - 1) / kNB;
  for (int J = 0; J < nJ; J++) {
    const int J0 = J * kNB;
    const int jb = std::min(kNB, rk - J0);
    hipLaunchKernelGGL(k_potrf_diag<false>, dim3(1), dim3(256), 0, st, P, rk, J0, Dinv, info_dev,
                       (const CholJob*)nullptr);
    ISDF_HIP_LAUNCHED("k_potrf_diag");
    const int rest = rk - J0 - jb;
    if (rest > 0) {
      hipLaunchKernelGGL(k_potrf_panel<false>, dim3(cdiv(rest, PanelTile::BM)), dim3(PanelTile::NT), 0,
                         st, P, rk, J0, Dinv, info_dev, (const CholJob*)nullptr);
      ISDF_HIP_LAUNCHED("k_potrf_panel");
      const unsigned t = cdiv(rest, SyrkTile::BM);
      hipLaunchKernelGGL(k_potrf_syrk<false>, dim3(t, t), dim3(SyrkTile::NT), 0, st, P, rk, J0, info_dev,
                         (const CholJob*)nullptr);
      ISDF_HIP_LAUNCHED("k_potrf_syrk");
    }
  }
  return 0;
}

int launch_make_fg(const double* L, int rk, const double* Dinv, double* Dinv128, double* Ft,
                   double* Gt, hipStream_t st) {
  const unsigned nJ = cdiv(rk, kSB);
  hipLaunchKernelGGL(k_dinv128<false>, dim3(nJ), dim3(256), 0, st, L, rk, Dinv, Dinv128,
                     (const CholJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_dinv128");
  hipLaunchKernelGGL(k_make_fg<false>, dim3(nJ, nJ), dim3(256), 0, st, L, rk, Dinv128, Ft, Gt,
                     (const CholJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_make_fg");
  return 0;
}

int launch_dinv128(const double* L, int rk, const double* Dinv, double* Dinv128, hipStream_t st) {
  hipLaunchKernelGGL(k_dinv128<false>, dim3(cdiv(rk, kSB)), dim3(256), 0, st, L, rk, Dinv,
                     Dinv128, (const CholJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_dinv128");
  return 0;
}

// Y = L^{-1} by recursive doubling, then its transpose.  Y starts as the
// block diagonal of the 128-block inverses; at half-size s, each pair of
// s-blocks (Y11 at (t, t), Y22 at (t+s, t+s), L21 the factor's block between)
// gets its lower-left block Y21 = -Y22 (L21 Y11), so every level is two
// batched GEMMs and rk = 1024 takes three levels.
__global__ __launch_bounds__(256) void k_linv_init(const double* __restrict__ Dinv128, int n,
                                                   double* __restrict__ Y) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int r = (int)(e % n), c = (int)(e / n);
  const int J = r / kSB;
  Y[e] = (c / kSB == J) ? Dinv128[(long long)J * kSB * kSB + (c % kSB) * kSB + r % kSB] : 0.0;
}

// 32 x 32 tiles, B read k-contiguous: at rk 1024 the three levels are 64, 128
// and 256 tiles (16, 32 and 64 of 64 x 64), on a chain that holds few CUs' work.
typedef MfmaTile<32, 32, 16, 2, 2, false> LinvTile;

// STAGE 0: T_p = L21 Y11 (s1 x s, ld s); STAGE 1: Y21 = -Y22 T_p.  Pair p =
// blockIdx.z starts at t = 2 s p; its lower block has s1 = min(s, n - t - s) rows.
template <int STAGE>
__global__ __launch_bounds__(LinvTile::NT) void k_linv_gemm(const double* __restrict__ L,
                                                            double* __restrict__ Y,
                                                            double* __restrict__ Tb, int n,
                                                            int s) {
  __shared__ double lds[LinvTile::LDS_DOUBLES];
  const long long t = 2LL * s * blockIdx.z;
  if (t + s >= n) return;
  const int s1 = (int)std::min<long long>(s, n - t - s);
  const int m0 = blockIdx.x * LinvTile::BM, n0 = blockIdx.y * LinvTile::BN;
  if (m0 >= s1) return;
  const int mlim = std::min(LinvTile::BM, s1 - m0), nlim = std::min(LinvTile::BN, s - n0);
  double* Tp = Tb + (long long)blockIdx.z * s * s;
  typename LinvTile::Acc acc;
  LinvTile::zero(acc);
  if constexpr (STAGE == 0) {
    // A(m, k) = L(t+s+m0+m, t+k); B(k, c) = Y(t+k, t+n0+c).
    LinvTile::run(L + t * n + t + s + m0, n, Y + (t + n0) * n + t, n, s, mlim, nlim, lds, acc);
    LinvTile::epilogue(acc, [&](int m, int c, double v) {
      if (m < mlim && c < nlim) Tp[(long long)(n0 + c) * s + m0 + m] = v;
    });
  } else {
    // A(m, k) = Y(t+s+m0+m, t+s+k); B(k, c) = T_p(k, n0+c).
    LinvTile::run(Y + (t + s) * n + t + s + m0, n, Tp + (long long)n0 * s, s, s1, mlim, nlim, lds,
                  acc);
    LinvTile::epilogue(acc, [&](int m, int c, double v) {
      if (m < mlim && c < nlim) Y[(t + n0 + c) * n + t + s + m0 + m] = -v;
    });
  }
}

// X = Y^T (n x n, column-major), through 32 x 32 LDS tiles.
__global__ __launch_bounds__(256) void k_transpose_sq(const double* __restrict__ Y, int n,
                                                      double* __restrict__ X) {
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) {
    const int r = bx + tx, c = by + j;
    tile[j][tx] = (r < n && c < n) ? Y[(long long)c * n + r] : 0.0;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int r = by + tx, c = bx + j;
    if (r < n && c < n) X[(long long)c * n + r] = tile[tx][j];
  }
}

int launch_linv_transposed(const double* L, int n, const double* Dinv128, double* Y, double* Tb,
                           double* X, hipStream_t st) {
  hipLaunchKernelGGL(k_linv_init, dim3(cdiv((long long)n * n, 256)), dim3(256), 0, st, Dinv128, n,
                     Y);
  ISDF_HIP_LAUNCHED("k_linv_init");
  for (int s = kSB; s < n; s *= 2) {
    const unsigned np = cdiv(n, 2 * s);
    const dim3 grid(cdiv(s, LinvTile::BM), cdiv(s, LinvTile::BN), np);
    hipLaunchKernelGGL(k_linv_gemm<0>, grid, dim3(LinvTile::NT), 0, st, L, Y, Tb, n, s);
    hipLaunchKernelGGL(k_linv_gemm<1>, grid, dim3(LinvTile::NT), 0, st, L, Y, Tb, n, s);
    ISDF_HIP_LAUNCHED("k_linv_gemm");
  }
  hipLaunchKernelGGL(k_transpose_sq, dim3(cdiv(n, 32), cdiv(n, 32)), dim3(256), 0, st,
                     (const double*)Y, n, X);
  ISDF_HIP_LAUNCHED("k_transpose_sq");
  return 0;
}

int potrs_variants() { return 14; }

namespace {
// Whether k_potrs_dpp with BM-row strips and KB-deep chunks can take the solve.
bool dpp_solve_fits(long long ng, long long ldt, int rk, const double* T, const double* Ft,
                    const double* Gt, int bm, int kb) {
  return ng % bm == 0 && rk % kb == 0 && ldt % 2 == 0 &&
         ((uintptr_t)T | (uintptr_t)Ft | (uintptr_t)Gt) % 16 == 0;
}

// k_potrs_dpp shapes: variant -> (TM, TG, KB, waves per SIMD).
template <int TM, int TG, int KB, int WPE>
int launch_potrs_dpp(double* T, long long ng, long long ldt, int rk, const double* Ft,
                     const double* Gt, hipStream_t st) {
  constexpr int BM = 64 * TM, NT = 64 * (kSB / (16 * TG));
  if (!dpp_solve_fits(ng, ldt, rk, T, Ft, Gt, BM, KB))
    return set_error(1, "the vector-FMA solve needs whole row strips, rk a multiple of its k "
                        "chunk and 16-byte aligned rows");
  hipLaunchKernelGGL((k_potrs_dpp<TM, TG, KB, WPE>), dim3((unsigned)(ng / BM)), dim3(NT), 0, st,
                     T, ldt, rk, Ft, Gt);
  return 0;
}
}  // namespace

int launch_potrs(double* T, long long ng, long long ldt, int rk, const double* Ft, const double* Gt,
                 int variant, hipStream_t st, bool unit_lower) {
  if (unit_lower && ng != rk) return set_error(1, "unit_lower needs a square identity operand");
  // Default: the vector-FMA solve where its shape conditions hold (9 on large grids, 10 on
  // small ones: 7% faster than the MFMA tile at 96^3, 5% on the dual 96^3 workload, 2% at
  // 64^3, 17% at 32^3), else the MFMA tiles (6, the LDS-staged write-back, measured
  // 1.5-2.5% faster than 0 on the 64^3 and 96^3 workloads).
  if (variant < 0) {
    const bool big = ng >= 256LL * 512;
    if (!unit_lower && dpp_solve_fits(ng, ldt, rk, T, Ft, Gt, 128, big ? 8 : 16))
      variant = big ? 9 : 10;
    else
      variant = big ? 6 : 2;
  }
  if (variant >= 9 && unit_lower) return set_error(1, "the vector-FMA solve runs both passes");
  switch (variant) {
#define ISDF_POTRS_CASE(V)                                                                      \
  case V:                                                                                       \
    hipLaunchKernelGGL((k_potrs<PotrsTile##V, false>), dim3(cdiv(ng, PotrsTile##V::BM)),               \
                       dim3(PotrsTile##V::NT), 0, st, T, ng, ldt, rk, Ft, Gt, (int)unit_lower,  \
                       (const CholJob*)nullptr);                                                \
    break;
    ISDF_POTRS_CASE(0)
    ISDF_POTRS_CASE(1)
    ISDF_POTRS_CASE(2)
    ISDF_POTRS_CASE(3)
    ISDF_POTRS_CASE(4)
    ISDF_POTRS_CASE(5)
    case 6:
      hipLaunchKernelGGL((k_potrs<PotrsTile0, false, true>), dim3(cdiv(ng, PotrsTile0::BM)),
                         dim3(PotrsTile0::NT), 0, st, T, ng, ldt, rk, Ft, Gt, (int)unit_lower,
                         (const CholJob*)nullptr);
      break;
    ISDF_POTRS_CASE(7)
    ISDF_POTRS_CASE(8)
#undef ISDF_POTRS_CASE
    case 9:
      if (int rc = launch_potrs_dpp<2, 2, 8, 3>(T, ng, ldt, rk, Ft, Gt, st)) return rc;
      break;
    case 10:
      if (int rc = launch_potrs_dpp<2, 1, 16, 4>(T, ng, ldt, rk, Ft, Gt, st)) return rc;
      break;
    case 11:
      if (int rc = launch_potrs_dpp<2, 1, 8, 4>(T, ng, ldt, rk, Ft, Gt, st)) return rc;
      break;
    case 12:
      if (int rc = launch_potrs_dpp<2, 2, 16, 2>(T, ng, ldt, rk, Ft, Gt, st)) return rc;
      break;
    case 13:
      if (int rc = launch_potrs_dpp<4, 1, 8, 2>(T, ng, ldt, rk, Ft, Gt, st)) return rc;
      break;
    default:
      return set_error(1, "unknown solve tile variant");
  }
  ISDF_HIP_LAUNCHED("k_potrs");
  return 0;
}

int launch_trace_eps_batched(const CholJob* jobs, int nb, int max_n, hipStream_t st) {
  (void)max_n;
  hipLaunchKernelGGL(k_trace_eps<true>, dim3(1, 1, nb), dim3(256), 0, st, (double*)nullptr, 0,
                     (double*)nullptr, jobs);
  ISDF_HIP_LAUNCHED("k_trace_eps (batched)");
  return 0;
}

int launch_potrf_batched(const CholJob* jobs, int nb, int max_n, hipStream_t st) {
  const int nJ = (max_n + kNB - 1) / kNB;
  for (int J = 0; J < nJ; J++) {
    const int J0 = J * kNB;
    hipLaunchKernelGGL(k_potrf_diag<true>, dim3(1, 1, nb), dim3(256), 0, st, (double*)nullptr, 0, J0,
                       (double*)nullptr, (int*)nullptr, jobs);
    ISDF_HIP_LAUNCHED("k_potrf_diag (batched)");
    const int rest = max_n - J0 - kNB;
    if (rest > 0) {
      hipLaunchKernelGGL(k_potrf_panel<true>, dim3(cdiv(rest, PanelTile::BM), 1, nb),
                         dim3(PanelTile::NT), 0, st, (double*)nullptr, 0, J0,
                         (const double*)nullptr, (const int*)nullptr, jobs);
      ISDF_HIP_LAUNCHED("k_potrf_panel (batched)");
      const unsigned t = cdiv(rest, SyrkTile::BM);
      hipLaunchKernelGGL(k_potrf_syrk<true>, dim3(t, t, nb), dim3(SyrkTile::NT), 0, st,
                         (double*)nullptr, 0, J0, (const int*)nullptr, jobs);
      ISDF_HIP_LAUNCHED("k_potrf_syrk (batched)");
    }
  }
  return 0;
}

int launch_make_fg_batched(const CholJob* jobs, int nb, int max_n, hipStream_t st) {
  const unsigned nJ = cdiv(max_n, kSB);
  hipLaunchKernelGGL(k_dinv128<true>, dim3(nJ, 1, nb), dim3(256), 0, st, (const double*)nullptr, 0,
                     (const double*)nullptr, (double*)nullptr, jobs);
  ISDF_HIP_LAUNCHED("k_dinv128 (batched)");
  hipLaunchKernelGGL(k_make_fg<true>, dim3(nJ, nJ, nb), dim3(256), 0, st, (const double*)nullptr, 0,
                     (const double*)nullptr, (double*)nullptr, (double*)nullptr, jobs);
  ISDF_HIP_LAUNCHED("k_make_fg (batched)");
  return 0;
}

int launch_potrs_identity_batched(const CholJob* jobs, int nb, int max_n, hipStream_t st) {
  hipLaunchKernelGGL((k_potrs<PotrsTile5, true>), dim3(cdiv(max_n, PotrsTile5::BM), 1, nb),
                     dim3(PotrsTile5::NT), 0, st, (double*)nullptr, 0LL, 0LL, 0,
                     (const double*)nullptr, (const double*)nullptr, 1, jobs);
  ISDF_HIP_LAUNCHED("k_potrs (batched identity)");
  return 0;
}

int launch_xxt(const double* X, int n, double* W, hipStream_t st) {
  const unsigned t = cdiv(n, XxtTile::BM);
  hipLaunchKernelGGL(k_xxt<false>, dim3(t, t), dim3(XxtTile::NT), 0, st, X, n, W,
                     (const CholJob*)nullptr);
  ISDF_HIP_LAUNCHED("k_xxt");
  return 0;
}

int launch_xxt_batched(const CholJob* jobs, int nb, int max_n, hipStream_t st) {
  const unsigned t = cdiv(max_n, XxtTile::BM);
  hipLaunchKernelGGL(k_xxt<true>, dim3(t, t, nb), dim3(XxtTile::NT), 0, st,
                     (const double*)nullptr, 0, (double*)nullptr, jobs);
  ISDF_HIP_LAUNCHED("k_xxt (batched)");
  return 0;
}

}  // namespace isdf_hip
