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
  // LDS row strides (doubles), 16 mod 32.  The four lane groups of a wave read
  // four k rows at once; a stride of 128 bytes past a multiple of 256 puts
  // them on disjoint banks instead of the same ones.
  static constexpr int AS = BM % 32 == 0 ? BM + 16 : BM;
  static constexpr int BS = BN % 32 == 0 ? BN + 16 : BN;
  static constexpr int LDS_DOUBLES = 2 * KB * (AS + BS);
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
    double* As = lds + buf * KB * (AS + BS);
    double* Bs = As + KB * AS;
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < APT; q++) {
      const int e = t + q * NT;
      As[(e / BM) * AS + e % BM] = ra[q];  // As[kk][m]
    }
#pragma unroll
    for (int q = 0; q < BPT; q++) {
      const int e = t + q * NT;
      if constexpr (B_KMAJOR) {
        Bs[(e / BN) * BS + e % BN] = rb[q];  // Bs[kk][n]
      } else {
        Bs[(e % KB) * BS + e / KB] = rb[q];
      }
    }
  }

  __device__ static void compute(const double* lds, int buf, Acc& acc) {
    const double* As = lds + buf * KB * (AS + BS);
    const double* Bs = As + KB * AS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % WGM, wn = wave / WGM;
    const int lr = lane & 15, lk = lane >> 4;
    const double* Aw = As + wm * (TM * 16) + lr;
    const double* Bw = Bs + wn * (TN * 16) + lr;
#pragma unroll
    for (int kk = 0; kk < KB; kk += 4) {
      double a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; i++) a[i] = Aw[(kk + lk) * AS + 16 * i];
#pragma unroll
      for (int j = 0; j < TN; j++) b[j] = Bw[(kk + lk) * BS + 16 * j];
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

  // Calls f(m, n, value) for every accumulator element of this lane (m, n
  // relative to the workgroup tile origin).
  template <class F>
  __device__ static void epilogue(const Acc& acc, F f) {
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int r = 0; r < 4; r++) f(row(i, r), col(j), acc[i][j][r]);
  }
  template <class F>
  __device__ static void epilogue2(const Acc& a1, const Acc& a2, F f) {
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int r = 0; r < 4; r++) f(row(i, r), col(j), a1[i][j][r], a2[i][j][r]);
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

// The same staged GEMM on the vector FMA pipe (v_fma_f64), which sustains a
// higher FP64 rate than v_mfma_f64_16x16x4f64 on gfx950.  Lane l of wave
// (wm, wn) owns rows wm*64*TM + 64*i + l (i < TM) and the wave-uniform columns
// wn*TN + j (j < TN): A values come from LDS per lane, B values as LDS
// broadcasts, so each k step is TM*TN FMAs for TM + TN LDS reads.
template <int BM_, int BN_, int KB_, int WGM_, int WGN_>
struct ValuTile {
  static constexpr int BM = BM_, BN = BN_, KB = KB_, WGM = WGM_, WGN = WGN_;
  static constexpr int NT = 64 * WGM * WGN;
  static constexpr int TM = BM / WGM / 64;
  static constexpr int TN = BN / WGN;
  static constexpr int APT = BM * KB / NT;
  static constexpr int BPT = BN * KB / NT;
  static constexpr int LDS_DOUBLES = 2 * KB * (BM + BN);
  static_assert(TM >= 1 && TN % 2 == 0 && BM % (64 * WGM) == 0 && BN % WGN == 0, "tile shape");
  static_assert(BM * KB % NT == 0 && BN * KB % NT == 0, "chunk must split evenly");
  typedef MfmaTile<BM, BN, KB, WGM, 1> Stager;  // same global->LDS staging (B k-major)
  typedef double Acc[TM][TN];

  __device__ static void zero(Acc& acc) {
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) acc[i][j] = 0.0;
  }

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
      const int kk = e / BN, n = e % BN;
      const int k = k0 + kk;
      rb[q] = (k < K && n < nlim) ? B[(long long)k * ldb + n] : 0.0;
    }
  }

  __device__ static void store(double* lds, int buf, const double (&ra)[APT],
                               const double (&rb)[BPT]) {
    double* As = lds + buf * KB * (BM + BN);
    double* Bs = As + KB * BM;
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < APT; q++) As[t + q * NT] = ra[q];
#pragma unroll
    for (int q = 0; q < BPT; q++) Bs[t + q * NT] = rb[q];
  }

  __device__ static void compute(const double* lds, int buf, Acc& acc) {
    const double* As = lds + buf * KB * (BM + BN);
    const double* Bs = As + KB * BM;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % WGM, wn = wave / WGM;
    const double* Aw = As + wm * (TM * 64) + lane;
    const double* Bw = Bs + wn * TN;
#pragma unroll 2
    for (int kk = 0; kk < KB; kk++) {
      double a[TM];
#pragma unroll
      for (int i = 0; i < TM; i++) a[i] = Aw[kk * BM + 64 * i];
      // B as pairs: one 16-byte LDS broadcast feeds two columns.
#pragma unroll
      for (int j = 0; j < TN; j += 2) {
        const double2 b = *(const double2*)(Bw + kk * BN + j);
#pragma unroll
        for (int i = 0; i < TM; i++) acc[i][j] = __builtin_fma(a[i], b.x, acc[i][j]);
#pragma unroll
        for (int i = 0; i < TM; i++) acc[i][j + 1] = __builtin_fma(a[i], b.y, acc[i][j + 1]);
      }
    }
  }

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

  __device__ static int row(int i) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    return (wave % WGM) * (TM * 64) + 64 * i + lane;
  }
  __device__ static int col(int j) {
    const int wave = threadIdx.x >> 6;
    return (wave / WGM) * TN + j;
  }
  template <class F>
  __device__ static void epilogue(const Acc& acc, F f) {
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int i = 0; i < TM; i++) f(row(i), col(j), acc[i][j]);
  }
  template <class F>
  __device__ static void epilogue2(const Acc& a1, const Acc& a2, F f) {
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int i = 0; i < TM; i++) f(row(i), col(j), a1[i][j], a2[i][j]);
  }
};

// acc[j] += b(lane j of this lane's row of 16) * a, for j = 0..15: sixteen
// v_fmac_f64 whose first operand comes through DPP row_newbcast:j, so the
// sixteen values of a row of lanes act as sixteen wave-uniform operands.  The
// leading s_nop 1 gives the two wait states a DPP read needs after a VALU
// write of its register (b is normally an LDS load, but the compiler may
// move it with a VALU copy).  The memory clobber keeps loads from moving
// across the block, so callers place their operand reads where they want them.
__device__ __forceinline__ void dpp_fma16(double* acc, double b, double a) {
#define ISDF_DF(J) "v_fmac_f64_dpp %" #J ", %16, %17 row_newbcast:" #J " row_mask:0xf bank_mask:0xf\n"
  asm volatile("s_nop 1\n" ISDF_DF(0) ISDF_DF(1) ISDF_DF(2) ISDF_DF(3) ISDF_DF(4) ISDF_DF(5)
               ISDF_DF(6) ISDF_DF(7) ISDF_DF(8) ISDF_DF(9) ISDF_DF(10) ISDF_DF(11) ISDF_DF(12)
               ISDF_DF(13) ISDF_DF(14) ISDF_DF(15)
               : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(acc[4]),
                 "+v"(acc[5]), "+v"(acc[6]), "+v"(acc[7]), "+v"(acc[8]), "+v"(acc[9]),
                 "+v"(acc[10]), "+v"(acc[11]), "+v"(acc[12]), "+v"(acc[13]), "+v"(acc[14]),
                 "+v"(acc[15])
               : "v"(b), "v"(a)
               : "memory");
#undef ISDF_DF
}

// One workgroup's acc += A (BM x K) * B (K x 128) on the FP64 vector FMA pipe,
// with the B operand broadcast by DPP (dpp_fma16) and both operands staged
// into LDS by LDS DMA.  BM = 64 TM rows; the NW = 128 / (16 TG) waves split
// the 128 columns into groups of 16 TG, each wave taking every row: lane l
// holds rows 64 i + l (i < TM) and columns 16 TG w + 16 g + j (g < TG, j < 16)
// in acc[(i TG + g) 16 + j].  Per k step a lane reads TM A values and TG B
// values from LDS (B at column 16 g + l % 16, so each row of 16 lanes holds a
// group) for 16 TM TG FMAs: an LDS broadcast of B would return every column
// to every lane, a scalar load of it would make each use wait for all.
//
// Staging: global_load_lds_dwordx4 writes 64 lanes x 16 bytes = 128 doubles
// of LDS from a wave-uniform base, each lane reading its own 16 bytes from
// anywhere; a chunk of KB k values is KB BM / 128 copies of A (column
// pieces, or whole columns in pairs when BM = 64) and KB copies of B rows, no
// VGPRs or LDS stores spent on it.  The copies of chunk c + 1 run during the
// FMAs of chunk c (double buffer, one barrier per chunk).
//
// Layouts: A(m, k) at A[m + k lda], B(k, n) at B[k ldb + n].  Requires K a
// multiple of KB, all BM rows of A readable (no row tail), 16-byte aligned
// column / row starts (lda, ldb even, A and B 16-byte aligned), and every B
// row readable for 128 columns: columns past the caller's width feed only
// output columns it does not store.
template <int TM, int TG, int KB>
struct DppGemm {
  static constexpr int BM = 64 * TM, BN = 128, NW = BN / (16 * TG), NT = 64 * NW;
  static constexpr int ACOPY = KB * BM / 128, NCOPY = ACOPY + KB;  // 1 KB copies per chunk
  static constexpr int CHUNK = KB * (BM + BN);                       // doubles per buffer
  static constexpr int LDS_DOUBLES = 2 * CHUNK;
  static constexpr int NACC = TM * TG * 16;
  static_assert(TG >= 1 && BN % (16 * TG) == 0, "column groups");
  static_assert(KB * BM % 128 == 0 && NCOPY % NW == 0, "copies must split over the waves");
  typedef double Acc[NACC];

  __device__ static void zero(Acc& acc) {
#pragma unroll
    for (int e = 0; e < NACC; e++) acc[e] = 0.0;
  }

  // This wave's copies of chunk c (k values c KB ..) into buffer `buf`.
  __device__ static void issue(const double* __restrict__ A, long long lda,
                               const double* __restrict__ B, long long ldb, int c, double* buf,
                               int wave, int lane) {
    typedef __attribute__((address_space(3))) void* LdsPtr;
    typedef __attribute__((address_space(1))) void* GlobalPtr;
    double* As = buf;
    double* Bs = buf + KB * BM;
#pragma unroll
    for (int q = 0; q < NCOPY / NW; q++) {
      const int idx = wave + q * NW;
      const double* src;
      double* dst;
      if (idx < ACOPY) {
        // LDS doubles idx 128 .. idx 128 + 127 of As = [KB][BM]; lane l's pair
        // 2 l, 2 l + 1 lies in one column (BM even).
        const int e = idx * 128 + 2 * lane;
        src = A + (long long)(c * KB + e / BM) * lda + e % BM;
        dst = As + idx * 128;
      } else {
        const int kk = idx - ACOPY;
        src = B + (long long)(c * KB + kk) * ldb + 2 * lane;
        dst = Bs + kk * BN;
      }
      __builtin_amdgcn_global_load_lds((GlobalPtr)src, (LdsPtr)dst, 16, 0, 0);
    }
  }

  // The FMAs of one chunk in buffer `buf`; the operands of step kk + 1 are read
  // before the FMAs of step kk (dpp_fma16 clobbers memory, so they stay there).
  __device__ static void compute(const double* buf, int wave, int lane, Acc& acc) {
    const double* Aw = buf + lane;
    const double* Bw = buf + KB * BM + wave * (16 * TG) + (lane & 15);
    double a[2][TM], b[2][TG];
#pragma unroll
    for (int i = 0; i < TM; i++) a[0][i] = Aw[64 * i];
#pragma unroll
    for (int g = 0; g < TG; g++) b[0][g] = Bw[16 * g];
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const int x = kk & 1;
      if (kk + 1 < KB) {
#pragma unroll
        for (int i = 0; i < TM; i++) a[x ^ 1][i] = Aw[(kk + 1) * BM + 64 * i];
#pragma unroll
        for (int g = 0; g < TG; g++) b[x ^ 1][g] = Bw[(kk + 1) * BN + 16 * g];
      }
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int g = 0; g < TG; g++) dpp_fma16(acc + (i * TG + g) * 16, b[x][g], a[x][i]);
    }
  }

  // acc += A * B over K (a multiple of KB).  Starts with a barrier (the LDS
  // buffers may still be read by a previous call's last chunk); ends after the
  // last chunk's FMAs, every copy of A and B landed (so A may be overwritten
  // afterwards), but other waves possibly still reading LDS.
  __device__ static void run(const double* __restrict__ A, long long lda,
                             const double* __restrict__ B, long long ldb, int K, double* lds,
                             Acc& acc) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nch = K / KB;
    __syncthreads();
    if (nch > 0) issue(A, lda, B, ldb, 0, lds, wave, lane);
    for (int c = 0; c < nch; c++) {
      // Chunk c has landed for every wave, and every wave is done with the
      // buffer the next copies overwrite (it held chunk c - 1).
      __syncthreads();
      if (c + 1 < nch) issue(A, lda, B, ldb, c + 1, lds + ((c + 1) & 1) * CHUNK, wave, lane);
      compute(lds + (c & 1) * CHUNK, wave, lane, acc);
    }
  }

  __device__ static int row(int i) { return 64 * i + (threadIdx.x & 63); }
  __device__ static int col(int g, int j) { return (threadIdx.x >> 6) * (16 * TG) + 16 * g + j; }
};

// Vector-FMA tile with the B operand in scalar registers and no LDS.  NW
// waves; lane l of wave w owns rows 64 i + l (i < TM = BM / 64) and the
// columns w TN + j (j < TN = BN / NW), the same for every lane of the wave:
// the B values of a k step are wave-uniform, so they come through scalar loads
// (constant address space) and feed v_fmac_f64 as its scalar operand, and the
// A values of a lane come from global memory through vector loads (the waves
// of a workgroup read the same A columns, shared in L1).  Scalar loads return
// out of order, so a use waits for all of them: the loads of step k + 1 are
// issued only after the first FMA of step k has taken its B value (an empty
// asm ties their address to that FMA's result), and they land during the rest
// of step k.  A is prefetched two steps ahead (vector loads return in order).
// B: row k at B[k * ldb], B_KMAJOR layout; the TN values of a row that a wave
// reads must lie in readable memory even past nlim (their results are never
// stored), so the caller's B allocation extends BN doubles past its end.
template <int BM_, int BN_, int NW_>
struct ScalarBTile {
  static constexpr int BM = BM_, BN = BN_, NW = NW_;
  static constexpr int NT = 64 * NW;
  static constexpr int TM = BM / 64;
  static constexpr int TN = BN / NW;
  static constexpr int LDS_DOUBLES = 1;  // no LDS staging
  static constexpr int WGN = NW;
  static_assert(BM % 64 == 0 && BN % NW == 0, "tile shape");
  typedef double Acc[TM][TN];
  static_assert(TN % 8 == 0, "B rows go as vectors of 8 doubles");
  typedef double D8 __attribute__((ext_vector_type(8)));
  typedef const __attribute__((address_space(4))) D8* CPtr8;

  __device__ static void zero(Acc& acc) {
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) acc[i][j] = 0.0;
  }

  // The wave's TN values of row `off / ldb` of B (scalar loads).
  __device__ static void load_b(const double* __restrict__ Bw, long long off, D8 (&b)[TN / 8]) {
#pragma unroll
    for (int h = 0; h < TN / 8; h++) b[h] = *(CPtr8)(Bw + off + 8 * h);
  }

  // One k step on register set X: the FMAs of column 0, then (their result
  // gating the address) the loads of step k + 1 into set Y, then the other
  // columns while those loads land.  Steps past K re-read step K - 1 (valid
  // memory; never used).
  template <int X, int Y>
  __device__ static void step(const double* __restrict__ A, const int (&ro)[TM], long long lda,
                              const double* __restrict__ Bw, long long ldb, int k, int K,
                              double (&a)[2][TM], D8 (&b)[2][TN / 8], Acc& acc) {
#pragma unroll
    for (int i = 0; i < TM; i++) acc[i][0] = __builtin_fma(a[X][i], b[X][0][0], acc[i][0]);
    long long kn = k + 1 < K ? k + 1 : K - 1;
    asm volatile("" : "+s"(kn) : "v"(acc[0][0]));
    // Volatile: left alone, the compiler sinks these loads to their use in
    // the next step and the two register sets collapse into one.
#pragma unroll
    for (int h = 0; h < TN / 8; h++) b[Y][h] = *(volatile CPtr8)(Bw + kn * ldb + 8 * h);
#pragma unroll
    for (int i = 0; i < TM; i++) a[Y][i] = A[kn * lda + ro[i]];
#pragma unroll
    for (int j = 1; j < TN; j++)
#pragma unroll
      for (int i = 0; i < TM; i++) acc[i][j] = __builtin_fma(a[X][i], b[X][j / 8][j % 8], acc[i][j]);
  }

  // acc += A[0:mlim, 0:K] * B[0:K, 0:nlim].  Ends with a workgroup barrier.
  // Rows at or past mlim read row mlim - 1 (their results are never stored).
  __device__ static void run(const double* __restrict__ A, long long lda,
                             const double* __restrict__ B, long long ldb, int K, int mlim,
                             int nlim, double* lds, Acc& acc) {
    (void)lds;
    (void)nlim;
    if (K > 0) {
      const int lane = threadIdx.x & 63;
      const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
      int ro[TM];  // the lane's rows, clamped into the strip
#pragma unroll
      for (int i = 0; i < TM; i++) ro[i] = min(lane + 64 * i, mlim - 1);
      const double* Bw = B + wave * TN;
      double a[2][TM];
      D8 b[2][TN / 8];
#pragma unroll
      for (int h = 0; h < TN / 8; h++) b[0][h] = *(CPtr8)(Bw + 8 * h);
#pragma unroll
      for (int i = 0; i < TM; i++) a[0][i] = A[ro[i]];
      for (int k = 0; k < K; k += 2) {
        step<0, 1>(A, ro, lda, Bw, ldb, k, K, a, b, acc);
        if (k + 1 >= K) break;
        step<1, 0>(A, ro, lda, Bw, ldb, k + 1, K, a, b, acc);
      }
    }
    __syncthreads();
  }

  __device__ static int row(int i) { return 64 * i + (threadIdx.x & 63); }
  __device__ static int col(int j) { return (threadIdx.x >> 6) * TN + j; }
  template <class F>
  __device__ static void epilogue(const Acc& acc, F f) {
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int i = 0; i < TM; i++) f(row(i), col(j), acc[i][j]);
  }
};

}  // namespace isdf_hip
