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
#include <type_traits>

#include "isdf_internal.h"
#include "mfma_gemm.h"

namespace isdf_hip {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));

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

// One GEMM operand of LpTile: ROWS rows of a 32-wide k chunk, staged from HBM
// into LDS by NT threads in 16-byte pieces.
//   FP32 operand: X(r, k) = ((const float*)X)[r*ld + k].  In the FP32 mode it
//     lands in LDS as FP32 rows of stride RS32; in the BF16X3 mode each value
//     is split (split3) into the three bf16 planes on the way.
//   Pre-split operand (BF16X3 only): the three planes interleaved per chunk,
//     plane p of X(r, k) at ((const __bf16*)X)[r*ld + (k/32)*96 + p*32 + k%32]
//     (ld = 3 K), copied to the planes unchanged: no conversion work in the loop.
// LDS planes: 64-byte rows of four 16-byte chunks, chunk c of row r stored at
// chunk c ^ ((r >> 2) & 3).  (Measured on gfx950, the fragment reads of this
// layout take two LDS cycles per ds_read_b128 instead of one; LpDmaTile's
// 128-byte rows are conflict-free.)
template <int ROWS, int NT, bool BF16, bool PRE>
struct LpOperand {
  static_assert(BF16 || !PRE, "pre-split operands are bf16");
  static constexpr int RS32 = 36;  // FP32 LDS row stride (floats): conflict-free reads
  static constexpr int RS16 = 32;  // bf16 plane row (elements)
  static constexpr int PIECES = PRE ? ROWS * 12 : ROWS * 8;
  static constexpr int PT = PIECES / NT;
  static constexpr int LDS_BYTES = BF16 ? 3 * ROWS * RS16 * 2 : ROWS * RS32 * 4;
  static_assert(PT >= 1 && PIECES % NT == 0, "staging split");
  typedef int i4 __attribute__((ext_vector_type(4)));
  typedef typename std::conditional<PRE, i4, f4>::type Reg;

  __device__ static int swz(int r, int c8) {
    return r * RS16 + 8 * ((c8 >> 1) ^ ((r >> 2) & 3)) + 4 * (c8 & 1);
  }

  __device__ static void load(const void* __restrict__ X, long long ld, int K, int k0, int lim,
                              Reg (&reg)[PT]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < PT; q++) {
      const int e = t + q * NT;
      if constexpr (PRE) {
        // Plane-major: 16 consecutive lanes store four whole 64-byte rows of
        // one plane (all 64 banks once) rather than three planes of one row.
        const int p = e / (ROWS * 4), r = (e >> 2) % ROWS, c16 = e & 3;
        reg[q] = r < lim ? *(const i4*)((const __bf16*)X + (long long)r * ld + (k0 / 32) * 96 +
                                        32 * p + 8 * c16)
                         : i4{0, 0, 0, 0};
      } else {
        const int r = e >> 3, k = k0 + 4 * (e & 7);
        reg[q] = (r < lim && k < K) ? *(const f4*)((const float*)X + (long long)r * ld + k)
                                    : f4{0.f, 0.f, 0.f, 0.f};
      }
    }
  }

  __device__ static void store(char* base, const Reg (&reg)[PT]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < PT; q++) {
      const int e = t + q * NT;
      if constexpr (PRE) {
        const int p = e / (ROWS * 4), r = (e >> 2) % ROWS, c16 = e & 3;
        __bf16* P = (__bf16*)base + p * ROWS * RS16;
        *(i4*)(P + r * RS16 + 8 * (c16 ^ ((r >> 2) & 3))) = reg[q];
      } else if constexpr (BF16) {
        __bf16* P = (__bf16*)base;
        const int o = swz(e >> 3, e & 7);
        bf4 h, m, l;
        split3(reg[q], h, m, l);
        *(bf4*)(P + o) = h;
        *(bf4*)(P + ROWS * RS16 + o) = m;
        *(bf4*)(P + 2 * ROWS * RS16 + o) = l;
      } else {
        *(f4*)((float*)base + (e >> 3) * RS32 + 4 * (e & 7)) = reg[q];
      }
    }
  }
};

// C[m][n] = sum_k A[m][k] B[n][k] over a BM x BN workgroup tile, K streamed in
// chunks of 32 through double-buffered LDS.  Operand formats as LpOperand
// (A_PRE / B_PRE: pre-split bf16 planes); K is a multiple of 32, FP32 operand
// rows are 16-byte aligned.  Rows m >= mlim and n >= nlim read as zero.
//
// 16x16 MFMA output layout (FP32 accumulators, measured on gfx950): lane l
// holds D[4*(l/16) + r][l % 16] in register r.
template <int BM_, int BN_, int WGM_, int WGN_, bool BF16_, bool A_PRE_ = false,
          bool B_PRE_ = false>
struct LpTile {
  static constexpr int BM = BM_, BN = BN_, WGM = WGM_, WGN = WGN_;
  static constexpr bool BF16 = BF16_, A_PRE = A_PRE_, B_PRE = B_PRE_;
  static constexpr int NT = 64 * WGM * WGN;
  static constexpr int TM = BM / WGM / 16;
  static constexpr int TN = BN / WGN / 16;
  static constexpr int KB = 32;
  typedef LpOperand<BM, NT, BF16, A_PRE> OpA;
  typedef LpOperand<BN, NT, BF16, B_PRE> OpB;
  static constexpr int RS32 = OpA::RS32, RS16 = OpA::RS16;
  static constexpr int BUF_BYTES = OpA::LDS_BYTES + OpB::LDS_BYTES;
  static constexpr int LDS_BYTES = 2 * BUF_BYTES;
  static_assert(TM >= 1 && TN >= 1, "tile too small for the wave grid");
  static_assert(LDS_BYTES <= 160 * 1024, "LDS");

  typedef f4 Acc[TM][TN];

  __device__ static void zero(Acc& acc) {
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  }

  __device__ static void compute(const char* lds, int buf, Acc& acc) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % WGM, wn = wave / WGM;
    const int lr = lane & 15, lk = lane >> 4;
    if constexpr (!BF16) {
      const float* As = (const float*)(lds + buf * BUF_BYTES);
      const float* Bs = (const float*)(lds + buf * BUF_BYTES + OpA::LDS_BYTES);
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
      const __bf16* Bp = (const __bf16*)(lds + buf * BUF_BYTES + OpA::LDS_BYTES);
      // Rows wm*TM*16 + 16 i + lr: (row >> 2) & 3 == (lr >> 2) & 3.
      const int co = lr * RS16 + 8 * (lk ^ ((lr >> 2) & 3));
      const __bf16* Aw = Ap + wm * TM * 16 * RS16 + co;
      const __bf16* Bw = Bp + wn * TN * 16 * RS16 + co;
      bf8 a[3][TM], b[3][TN];
#pragma unroll
      for (int p = 0; p < 3; p++) {
#pragma unroll
        for (int i = 0; i < TM; i++) a[p][i] = *(const bf8*)(Aw + (p * BM + 16 * i) * RS16);
#pragma unroll
        for (int j = 0; j < TN; j++) b[p][j] = *(const bf8*)(Bw + (p * BN + 16 * j) * RS16);
      }
      // Six partial products, smallest first; each pass issues TM*TN independent MFMAs.
      constexpr int PA[6] = {0, 2, 1, 0, 1, 0};
      constexpr int PB[6] = {2, 0, 1, 1, 0, 0};
#pragma unroll
      for (int p = 0; p < 6; p++)
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[PA[p]][i], b[PB[p]][j],
                                                                acc[i][j], 0, 0, 0);
    }
  }

  // acc += A[0:mlim, 0:K] B[0:nlim, 0:K]^T.  Ends with a workgroup barrier.
  __device__ static void run(const void* __restrict__ A, long long lda,
                             const void* __restrict__ B, long long ldb, int K, int mlim,
                             int nlim, char* lds, Acc& acc) {
    typename OpA::Reg ra[OpA::PT];
    typename OpB::Reg rb[OpB::PT];
    const int nch = (K + KB - 1) / KB;
    if (nch == 0) return;
    OpA::load(A, lda, K, 0, mlim, ra);
    OpB::load(B, ldb, K, 0, nlim, rb);
    OpA::store(lds, ra);
    OpB::store(lds + OpA::LDS_BYTES, rb);
    __syncthreads();
    for (int c = 0; c < nch; c++) {
      const bool more = c + 1 < nch;
      if (more) {
        OpA::load(A, lda, K, (c + 1) * KB, mlim, ra);
        OpB::load(B, ldb, K, (c + 1) * KB, nlim, rb);
      }
      compute(lds, c & 1, acc);
      if (more) {
        char* nb = lds + ((c + 1) & 1) * BUF_BYTES;
        OpA::store(nb, ra);
        OpB::store(nb + OpA::LDS_BYTES, rb);
      }
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

// BF16X3 tile with both operands FP32 in HBM, moved HBM -> LDS by
// direct-to-LDS loads (global_load_lds_dwordx4, 16 bytes per lane, lane L of
// a wave landing at byte 16 L of a 1 KB LDS block), into a ring of STAGES
// LDS buffers (STAGES-1 k chunks in flight while one is multiplied).  With no
// staging registers a 256 x 256 tile fits in 256 VGPRs (two waves per SIMD),
// which halves the W traffic per flop of the 256 x 128 LpTile; measured on the
// 96^3 workload, a third stage gains nothing (the loads are not the limit).
// Each operand row of a chunk (32 FP32) is one 128-byte LDS row
// whose 16-byte piece c sits at position c ^ sw(r); the fragment reads (16
// lanes on 16 consecutive rows, one ds_read_b128 each) measured free of bank
// conflicts in that layout on gfx950.  Fragments are split into their bf16
// parts h/m/l (the arithmetic of split3) right after the read, so the products
// and their order match LpTile's BF16X3 mode bit for bit.
//   Operand rows are FP32 with row stride ld (a multiple of 4), K a multiple of
//   32; rows >= mlim / nlim are read as row mlim-1 / nlim-1 (their outputs are
//   not stored), so every load stays inside the operands.
//
// B_SPLIT: the B operand comes pre-split from HBM (the three bf16 parts
// interleaved per 32-wide k chunk, row stride 3 K, as LpOperand's pre-split
// rows) and needs no conversion.  In LDS, row r's h and m parts share one
// 128-byte row (16-byte piece q at q ^ (r & 7)), and the l parts of rows 2R
// and 2R+1 share row R of a second region (piece 4 (r & 1) + c at that
// ^ (R & 7)); both measured conflict-free for the fragment reads.
template <int BM_, int BN_, int WGM_, int WGN_, int STAGES_ = 2, bool B_SPLIT_ = false>
struct LpDmaTile {
  static constexpr int BM = BM_, BN = BN_, WGM = WGM_, WGN = WGN_, STAGES = STAGES_;
  static constexpr bool BF16 = true, A_PRE = false, B_PRE = B_SPLIT_, B_SPLIT = B_SPLIT_;
  static constexpr int NW = WGM * WGN, NT = 64 * NW;
  static constexpr int TM = BM / WGM / 16, TN = BN / WGN / 16;
  static constexpr int KB = 32, ROWB = 128;
  static_assert(STAGES >= 2, "stages");
  static constexpr int A_BYTES = BM * ROWB;
  static constexpr int B_HM_BYTES = BN * ROWB;  // FP32 rows, or the h|m rows
  static constexpr int STAGE_BYTES = A_BYTES + B_HM_BYTES + (B_SPLIT ? BN / 2 * ROWB : 0);
  static constexpr int LDS_BYTES = STAGES * STAGE_BYTES;
  static constexpr int A_PW = BM / 8 / NW, B_PW = BN / 8 / NW;  // 1 KB blocks per wave
  static constexpr int BL_PW = B_SPLIT ? BN / 16 / NW : 0;      // l-part blocks per wave
  static constexpr int NDMA = A_PW + B_PW + BL_PW;              // loads per wave per chunk
  static_assert(A_PW * 8 * NW == BM && B_PW * 8 * NW == BN, "DMA blocks per wave");
  static_assert(!B_SPLIT || BL_PW * 16 * NW == BN, "l-part blocks per wave");
  static_assert(LDS_BYTES <= 160 * 1024, "LDS");
  static_assert(NDMA * 2 <= 63 && STAGES <= 4, "vmcnt range");
  typedef f4 Acc[TM][TN];
  typedef __attribute__((address_space(3))) void* LdsPtr;

  __device__ static int sw(int r) { return ((r & 3) << 1) | ((r >> 2) & 1); }

  __device__ static void zero(Acc& acc) {
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  }

  // Chunk at k0 of both operands -> stage buffer `stage` (NDMA loads per wave).
  __device__ static void issue(const float* __restrict__ A, long long lda,
                               const float* __restrict__ B, long long ldb, int k0, int mlim,
                               int nlim, char* stage) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, rl = lane >> 3;
    const int c4 = 4 * ((lane & 7) ^ sw(rl));
#pragma unroll
    for (int j = 0; j < A_PW; j++) {
      const int blk = wave + NW * j, r = std::min(blk * 8 + rl, mlim - 1);
      __builtin_amdgcn_global_load_lds((const void*)(A + (long long)r * lda + k0 + c4),
                                       (LdsPtr)(stage + blk * 1024), 16, 0, 0);
    }
    if constexpr (!B_SPLIT) {
#pragma unroll
      for (int j = 0; j < B_PW; j++) {
        const int blk = wave + NW * j, r = std::min(blk * 8 + rl, nlim - 1);
        __builtin_amdgcn_global_load_lds((const void*)(B + (long long)r * ldb + k0 + c4),
                                         (LdsPtr)(stage + A_BYTES + blk * 1024), 16, 0, 0);
      }
    } else {
      // B is bf16 here (ldb in bf16 elements); chunk k0 of a row starts at
      // element (k0 / 32) * 96: h at +0, m at +32, l at +64.
      const __bf16* Bh = (const __bf16*)(const void*)B + (k0 / 32) * 96;
#pragma unroll
      for (int j = 0; j < B_PW; j++) {
        const int blk = wave + NW * j, row = blk * 8 + rl;
        const int q = (lane & 7) ^ (row & 7), r = std::min(row, nlim - 1);
        __builtin_amdgcn_global_load_lds((const void*)(Bh + (long long)r * ldb + 8 * q),
                                         (LdsPtr)(stage + A_BYTES + blk * 1024), 16, 0, 0);
      }
#pragma unroll
      for (int j = 0; j < BL_PW; j++) {
        const int blk = wave + NW * j, R = blk * 8 + rl;
        const int q = (lane & 7) ^ (R & 7);
        const int r = std::min(2 * R + (q >> 2), nlim - 1);
        __builtin_amdgcn_global_load_lds(
            (const void*)(Bh + (long long)r * ldb + 64 + 8 * (q & 3)),
            (LdsPtr)(stage + A_BYTES + B_HM_BYTES + blk * 1024), 16, 0, 0);
      }
    }
  }

  // This wave's loads except the newest N have landed, its LDS reads are
  // done, and so are every other wave's (the barrier).  The asm is also a
  // compiler memory barrier, so no LDS access moves across it.
  template <int N>
  __device__ static void wait_barrier() {
    asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(N) : "memory");
  }

  // split3 on a pair: one two-source bf16 conversion per part, the parts
  // widened back to FP32 by shift / mask of the packed pair.
  __device__ static void split2(float x0, float x1, unsigned& h, unsigned& m, unsigned& l) {
    h = __builtin_bit_cast(unsigned, __builtin_convertvector(f2{x0, x1}, bf2));
    const float r0 = x0 - __uint_as_float(h << 16), r1 = x1 - __uint_as_float(h & 0xffff0000u);
    m = __builtin_bit_cast(unsigned, __builtin_convertvector(f2{r0, r1}, bf2));
    l = __builtin_bit_cast(unsigned, __builtin_convertvector(
                                         f2{r0 - __uint_as_float(m << 16),
                                            r1 - __uint_as_float(m & 0xffff0000u)},
                                         bf2));
  }

  __device__ static void split8(f4 x0, f4 x1, bf8& h, bf8& m, bf8& l) {
    unsigned h0, h1, h2, h3, m0, m1, m2, m3, l0, l1, l2, l3;
    split2(x0[0], x0[1], h0, m0, l0);
    split2(x0[2], x0[3], h1, m1, l1);
    split2(x1[0], x1[1], h2, m2, l2);
    split2(x1[2], x1[3], h3, m3, l3);
    const u4 hh = {h0, h1, h2, h3}, mm = {m0, m1, m2, m3}, ll = {l0, l1, l2, l3};
    h = __builtin_bit_cast(bf8, hh);
    m = __builtin_bit_cast(bf8, mm);
    l = __builtin_bit_cast(bf8, ll);
  }

  __device__ static void compute(const char* stage, Acc& acc) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % WGM, wn = wave / WGM;
    const int lr = lane & 15, lk = lane >> 4, s = sw(lr);
    const int p0 = 16 * ((2 * lk) ^ s), p1 = 16 * ((2 * lk + 1) ^ s);
    const char* Aw = stage + (wm * TM * 16 + lr) * ROWB;
    const char* Bw = stage + A_BYTES + (wn * TN * 16 + lr) * ROWB;
    bf8 a[3][TM];
#pragma unroll
    for (int i = 0; i < TM; i++)
      split8(*(const f4*)(Aw + 16 * ROWB * i + p0), *(const f4*)(Aw + 16 * ROWB * i + p1),
             a[0][i], a[1][i], a[2][i]);
    // One B fragment at a time (its three parts live only for its column of
    // accumulators); each accumulator still sees the six products in LpTile's
    // order, so the result is the same bit for bit.
    constexpr int PA[6] = {0, 2, 1, 0, 1, 0};
    constexpr int PB[6] = {2, 0, 1, 1, 0, 0};
    // Pre-split B: row r = wn*TN*16 + 16 j + lr, so r & 7 = lr & 7, r & 1 =
    // lr & 1 and (r >> 1) & 7 = (lr >> 1) & 7.
    const int ph = 16 * (lk ^ (lr & 7)), pm = 16 * ((4 + lk) ^ (lr & 7));
    const int pl = 16 * ((((lr & 1) << 2) + lk) ^ ((lr >> 1) & 7));
    const char* Bl = stage + A_BYTES + B_HM_BYTES + ((wn * TN * 16 + lr) >> 1) * ROWB;
#pragma unroll
    for (int j = 0; j < TN; j++) {
      bf8 b[3];
      if constexpr (B_SPLIT) {
        b[0] = *(const bf8*)(Bw + 16 * ROWB * j + ph);
        b[1] = *(const bf8*)(Bw + 16 * ROWB * j + pm);
        b[2] = *(const bf8*)(Bl + 8 * ROWB * j + pl);
      } else {
        split8(*(const f4*)(Bw + 16 * ROWB * j + p0), *(const f4*)(Bw + 16 * ROWB * j + p1),
               b[0], b[1], b[2]);
      }
#pragma unroll
      for (int p = 0; p < 6; p++)
#pragma unroll
        for (int i = 0; i < TM; i++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[PA[p]][i], b[PB[p]], acc[i][j],
                                                              0, 0, 0);
    }
  }

  // acc += A[0:mlim, 0:K] B[0:nlim, 0:K]^T.  Ends with a workgroup barrier.
  __device__ static void run(const void* __restrict__ Av, long long lda,
                             const void* __restrict__ Bv, long long ldb, int K, int mlim,
                             int nlim, char* lds, Acc& acc) {
    const float* A = (const float*)Av;
    const float* B = (const float*)Bv;
    const int nch = K / KB;
    if (nch == 0) return;
    // Chunks 0 .. STAGES-2 in flight before the loop; iteration c waits for
    // chunk c, then refills the stage freed by chunk c-1 with chunk c+STAGES-1.
    constexpr int AHEAD = STAGES - 1;
#pragma unroll
    for (int c = 0; c < AHEAD; c++)
      if (c < nch) issue(A, lda, B, ldb, c * KB, mlim, nlim, lds + c * STAGE_BYTES);
    int cur = 0, nxt = AHEAD;  // stage of chunk c, stage of chunk c + AHEAD
    for (int c = 0; c < nch; c++) {
      // Loads issued after chunk c's: chunks c+1 .. min(c+AHEAD-1, nch-1).
      const int later = std::min(AHEAD - 1, nch - 1 - c);
      if constexpr (AHEAD >= 3) {
        if (later >= 2) wait_barrier<2 * NDMA>();
        else if (later == 1) wait_barrier<NDMA>();
        else wait_barrier<0>();
      } else if constexpr (AHEAD == 2) {
        if (later >= 1) wait_barrier<NDMA>();
        else wait_barrier<0>();
      } else {
        wait_barrier<0>();
      }
      if (c + AHEAD < nch) issue(A, lda, B, ldb, (c + AHEAD) * KB, mlim, nlim, lds + nxt * STAGE_BYTES);
      compute(lds + cur * STAGE_BYTES, acc);
      cur = cur == STAGES - 1 ? 0 : cur + 1;
      nxt = nxt == STAGES - 1 ? 0 : nxt + 1;
    }
    wait_barrier<0>();
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

// Writes operand element (r, k) of a row-major operand with kp (multiple of
// 32) columns: FP32 at out[r*kp + k], or, SPLIT, its three bf16 parts in the
// pre-split layout of LpOperand (row stride 3 kp).
template <bool SPLIT>
__device__ __forceinline__ void put_operand(void* out, int kp, long long r, int k, float v) {
  if constexpr (SPLIT) {
    __bf16* o = (__bf16*)out + r * 3 * kp + (k >> 5) * 96 + (k & 31);
    const __bf16 h = (__bf16)v;
    const float r1 = v - (float)h;
    const __bf16 m = (__bf16)r1;
    o[0] = h;
    o[32] = m;
    o[64] = (__bf16)(r1 - (float)m);
  } else {
    ((float*)out)[r * kp + k] = v;
  }
}

// Packs columns [c0, c0+K) of the FP64 column-major psi (ld) into operand rows
// g (FP32 or pre-split), zero for K <= k < kp.  64 grid points x 32 states per
// block (block bx, by of the grid cdiv(ng, 64) x cdiv(kp, 32)), transposed
// through LDS so that both sides are coalesced.
template <bool SPLIT>
__device__ __forceinline__ void pack_rows_block(const double* __restrict__ psi, long long ld,
                                                int ng, int c0, int K, int kp,
                                                void* __restrict__ out, int bx, int by) {
  __shared__ float tile[32][65];
  const long long g0 = (long long)bx * 64;
  const int k0 = by * 32;
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int kk = (t >> 6) + 4 * q, gg = t & 63;
    const long long g = g0 + gg;
    const int k = k0 + kk;
    tile[kk][gg] = (g < ng && k < K) ? (float)psi[(long long)(c0 + k) * ld + g] : 0.f;
  }
  __syncthreads();
  if constexpr (SPLIT) {
    // Thread t splits 8 consecutive k of row t/4 and writes the three 16-byte
    // plane pieces of the row's 192-byte chunk.
    const int gg = t >> 2, k8 = (t & 3) * 8;
    const long long g = g0 + gg;
    if (g < ng) {
      bf8 h, m, l;
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const float v = tile[k8 + u][gg];
        const __bf16 a = (__bf16)v;
        const float r1 = v - (float)a;
        const __bf16 b = (__bf16)r1;
        h[u] = a;
        m[u] = b;
        l[u] = (__bf16)(r1 - (float)b);
      }
      __bf16* o = (__bf16*)out + g * 3 * kp + (k0 / 32) * 96 + k8;
      *(bf8*)o = h;
      *(bf8*)(o + 32) = m;
      *(bf8*)(o + 64) = l;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int gg = (t >> 5) + 8 * q, kk = t & 31;
      const long long g = g0 + gg;
      if (g < ng) put_operand<SPLIT>(out, kp, g, k0 + kk, tile[kk][gg]);
    }
  }
}

template <bool SPLIT>
__global__ __launch_bounds__(256) void k_pack_rows(const double* __restrict__ psi, long long ld,
                                                   int ng, int c0, int K, int kp,
                                                   void* __restrict__ out) {
  pack_rows_block<SPLIT>(psi, ld, ng, c0, K, kp, out, blockIdx.x, blockIdx.y);
}

// Operand row nu = psi(piv[nu], c0 + k) (FP32 or pre-split), zero for k >= K;
// element t = nu kp + k.
template <bool SPLIT>
__device__ __forceinline__ void gather_rows_elem(const double* __restrict__ psi, long long ld,
                                                 const int* __restrict__ piv, int rk, int c0,
                                                 int K, int kp, void* __restrict__ out,
                                                 long long t) {
  if (t >= (long long)rk * kp) return;
  const int k = (int)(t % kp), nu = (int)(t / kp);
  put_operand<SPLIT>(out, kp, nu, k,
                     k < K ? (float)psi[(long long)(c0 + k) * ld + piv[nu]] : 0.f);
}

template <bool SPLIT>
__global__ __launch_bounds__(256) void k_gather_rows(const double* __restrict__ psi, long long ld,
                                                     const int* __restrict__ piv, int rk, int c0,
                                                     int K, int kp, void* __restrict__ out) {
  gather_rows_elem<SPLIT>(psi, ld, piv, rk, c0, K, kp, out,
                          (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// The first blocks of a batched launch's jobs (field read through `start`) into
// LDS, once per workgroup.
template <class Job, class Start>
__device__ __forceinline__ void load_job_starts(const Job* __restrict__ jobs, int nb, int* s0,
                                                Start start) {
  for (int j = threadIdx.x; j < nb; j += blockDim.x) s0[j] = start(jobs[j]);
  __syncthreads();
}

// The job of block b: the last job j with s0[j] <= b.
__device__ __forceinline__ int job_of(const int* s0, int nb, int b) {
  int j = 0;
  while (j + 1 < nb && s0[j + 1] <= b) j++;
  return j;
}

// Every instance's operand rows in two launches (the per-instance
// launch_pack_psi / launch_gather_rows): the blocks of job j are
// pack_b0 .. pack_b0 + cdiv(ng, 64) cdiv(kp, 32) - 1 of the pack block list and
// gather_b0 .. gather_b0 + cdiv(rk kp, 256) - 1 of the gather grid.  The pack
// runs beside the K-means, so it walks its block list with a bounded grid and
// leaves room on every CU for the K-means kernels.
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_pack_rows_batched(const FrontJob* __restrict__ jobs,
                                                           int nb, int nblocks) {
  __shared__ int s0[kMaxFrontJobs];
  load_job_starts(jobs, nb, s0, [](const FrontJob& f) { return f.pack_b0; });
  for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const FrontJob& f = jobs[job_of(s0, nb, b)];
    const int lb = b - f.pack_b0, nbx = (f.ng + 63) / 64;
    pack_rows_block<SPLIT>(f.psi, f.ld, f.ng, f.c0, f.K, f.kp, f.A, lb % nbx, lb / nbx);
    __syncthreads();  // the block's LDS tile is rewritten by the next one
  }
}

template <bool SPLIT>
__global__ __launch_bounds__(256) void k_gather_rows_batched(const FrontJob* __restrict__ jobs,
                                                             int nb) {
  __shared__ int s0[kMaxFrontJobs];
  load_job_starts(jobs, nb, s0, [](const FrontJob& f) { return f.gather_b0; });
  const int b = blockIdx.x;
  const FrontJob& f = jobs[job_of(s0, nb, b)];
  gather_rows_elem<SPLIT>(f.psi, f.ld, f.piv, f.rk, f.c0, f.K, f.kp, f.B,
                          (long long)(b - f.gather_b0) * 256 + threadIdx.x);
}

// Zt[g][nu] = P1(g, nu) * P2(g, nu) (P2 = P1 when the pair sets agree), for
// nu < ldz (columns rk..ldz come out zero, as the Theta GEMM's k padding).
// A1/A2 (grid rows) and B1/B2 (pivot rows) in the Tile's operand formats, row
// strides lda1/lda2 (elements of the format).
// ZSPLIT: Zt is written pre-split (the three bf16 parts per 32-wide chunk,
// row stride 3 ldz) for a Theta tile that reads it without conversion.
template <class Tile, bool DUAL, bool ZSPLIT = false, bool PERSIST = false>
__global__ __launch_bounds__(Tile::NT, 1) void k_zct_lp(int ng, int rk, int ldz,
                                                        const void* __restrict__ A1,
                                                        const void* __restrict__ A2, int K1,
                                                        int K2, long long lda1, long long lda2,
                                                        const void* __restrict__ B1,
                                                        const void* __restrict__ B2,
                                                        float* __restrict__ Zt) {
  __shared__ __attribute__((aligned(16))) char lds[Tile::LDS_BYTES];
  const int nstrip = (ng + Tile::BM - 1) / Tile::BM;
  const int ntile = (ldz + Tile::BN - 1) / Tile::BN;
  // One tile per workgroup, or (PERSIST: a smaller grid, a multiple of 8)
  // each workgroup walks tiles tb, tb + gridDim.x, ... on the same XCD; the
  // one-tile form is its own instantiation (the loop costs registers).
  for (int tb = blockIdx.x; tb < nstrip * ntile; tb += gridDim.x) {
  int s, ct;
  xcd_map(tb, ntile, nstrip, &s, &ct);
  if (s >= nstrip) continue;
  const long long m0 = (long long)s * Tile::BM;
  const int n0 = ct * Tile::BN;
  const int mlim = (int)std::min<long long>(Tile::BM, ng - m0);
  const int nlim = std::min(Tile::BN, rk - n0);   // operand rows that exist
  const int wlim = std::min(Tile::BN, ldz - n0);  // columns written (zeros past rk)
  constexpr int ESA = Tile::A_PRE ? 2 : 4, ESB = Tile::B_PRE ? 2 : 4;  // element bytes
  typename Tile::Acc acc1, acc2;
  Tile::zero(acc1);
  Tile::run((const char*)A1 + m0 * lda1 * ESA, lda1, (const char*)B1 + (long long)n0 * lda1 * ESB,
            lda1, K1, mlim, nlim, lds, acc1);
  if constexpr (DUAL) {
    Tile::zero(acc2);
    Tile::run((const char*)A2 + m0 * lda2 * ESA, lda2,
              (const char*)B2 + (long long)n0 * lda2 * ESB, lda2, K2, mlim, nlim, lds, acc2);
  }
  if constexpr (ZSPLIT) {
    // Pre-split Zt through LDS (the staging buffers are free after run()):
    // each wave stages 32 of its columns at a time for its 16 TM rows, then
    // every lane splits 8 consecutive columns of one row and stores the three
    // parts as 16-byte pieces of the row's 192-byte chunk.
    static_assert(Tile::NW * 16 * Tile::TM * 36 * 4 <= Tile::LDS_BYTES, "epilogue staging");
    constexpr int WR = 16 * Tile::TM, WS = 36;  // wave rows, LDS row stride (floats)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float* Ws = (float*)(void*)lds + wave * WR * WS;
    const int wr0 = (wave % Tile::WGM) * WR, wc0 = (wave / Tile::WGM) * (16 * Tile::TN);
#pragma unroll
    for (int jp = 0; jp < Tile::TN; jp += 2) {
#pragma unroll
      for (int jj = 0; jj < 2; jj++) {
        const int n = Tile::col(jp + jj);
#pragma unroll
        for (int i = 0; i < Tile::TM; i++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const float v1 = acc1[i][jp + jj][r];
            const float v2 = DUAL ? acc2[i][jp + jj][r] : v1;
            // Columns past rk are zero (the tile clamps operand rows).
            Ws[(16 * i + 4 * (lane >> 4) + r) * WS + 16 * jj + (lane & 15)] =
                n < nlim ? v1 * v2 : 0.f;
          }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < WR * 4 / 64; q++) {
        const int item = lane + 64 * q, lr = item >> 2, g8 = item & 3;
        const int m = wr0 + lr, nc = wc0 + 16 * jp + 8 * g8;  // nc: tile column of the piece
        if (m < mlim && nc < wlim) {
          const f4 x0 = *(const f4*)(Ws + lr * WS + 8 * g8);
          const f4 x1 = *(const f4*)(Ws + lr * WS + 8 * g8 + 4);
          bf8 h, mm, l;
          LpDmaTile<256, 128, 4, 2>::split8(x0, x1, h, mm, l);
          const int nn = n0 + nc;
          __bf16* o = (__bf16*)(void*)Zt + (m0 + m) * 3LL * ldz + (nn >> 5) * 96 + (nn & 31);
          *(bf8*)o = h;
          *(bf8*)(o + 32) = mm;
          *(bf8*)(o + 64) = l;
        }
      }
      __syncthreads();
    }
  } else {
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
            // Columns past rk are zero (a tile that clamps operand rows
            // instead of zeroing them leaves other values in those
            // accumulators).
            Zm[(long long)m * ldz + n] = n < nlim ? v1 * v2 : 0.f;
          }
        }
    }
  }
  if constexpr (!PERSIST) break;
  }
}

// P(mu, nu) = Zt[piv[mu]][nu] (column-major rk x rk FP64),
// src/cpu/getbasis_scalapack.cpp:390-396; piv == NULL reads row mu (Zt holds
// the pivot rows only).
__global__ __launch_bounds__(256) void k_cct_lp(const float* __restrict__ Zt, int ldz,
                                                const int* __restrict__ piv, int rk,
                                                double* __restrict__ P) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)rk * rk) return;
  const int mu = (int)(t % rk), nu = (int)(t / rk);
  const long long row = piv ? piv[mu] : mu;
  P[t] = (double)Zt[row * ldz + nu];
}

// Operand row nu of W: W(k, nu) (FP32 or pre-split, ldw columns), zero for
// k >= rk (W column-major FP64).
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_pack_w(const double* __restrict__ W, int rk, int ldw,
                                                void* __restrict__ Wf) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)rk * ldw) return;
  const int k = (int)(t % ldw), nu = (int)(t / ldw);
  put_operand<SPLIT>(Wf, ldw, nu, k, k < rk ? (float)W[(long long)nu * rk + k] : 0.f);
}

// theta(g, nu) (=|+=) sum_k W(k, nu) Zt[g][k]: rows nu of the tile are the A
// operand (Wf, row stride ldwr elements of its format), grid points g the B
// operand (FP32 Zt, row stride ldk); each lane group writes 16 consecutive g
// of one column of the column-major theta.
template <class Tile, bool ADD>
__global__ __launch_bounds__(Tile::NT, 1) void k_theta_lp(int ng, int rk, int ldk,
                                                          const void* __restrict__ Wf,
                                                          long long ldwr,
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
  constexpr int ESA = Tile::A_PRE ? 2 : 4, ESB = Tile::B_PRE ? 2 : 4;
  const long long ldbz = Tile::B_PRE ? 3LL * ldk : ldk;  // Zt row stride (elements)
  typename Tile::Acc acc;
  Tile::zero(acc);
  Tile::run((const char*)Wf + (long long)v0 * ldwr * ESA, ldwr,
            (const char*)Zt + g0 * ldbz * ESB, ldbz, ldk, mlim, nlim, lds, acc);
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

// Grouped (persistent) forms of the two contractions: `gridDim.x` workgroups
// walk one tile list that spans every job; tile t goes to workgroup t % grid,
// and with grid a multiple of 8 and each job's list starting at a multiple of
// 8, a job's tiles meet the XCDs in the same pattern as in the single launch.
template <class Tile>
__global__ __launch_bounds__(Tile::NT, 1) void k_zct_lp_grouped(const LpGemmJob* __restrict__ jobs,
                                                                int njobs, long long total) {
  __shared__ __attribute__((aligned(16))) char lds[Tile::LDS_BYTES];
  constexpr int ES = Tile::A_PRE ? 2 : 4;
  for (long long t = blockIdx.x; t < total; t += gridDim.x) {
    int j = 0;
    while (j + 1 < njobs && jobs[j + 1].tile0 <= t) j++;
    const LpGemmJob J = jobs[j];
    const int nstrip = (J.m + Tile::BM - 1) / Tile::BM;
    const int ntile = (J.ldc + Tile::BN - 1) / Tile::BN;
    const long long local = t - J.tile0;
    if (local >= (long long)nstrip * ntile) continue;  // padding of the job's list
    int s, ct;
    xcd_map((int)local, ntile, nstrip, &s, &ct);
    const long long m0 = (long long)s * Tile::BM;
    const int n0 = ct * Tile::BN;
    const int mlim = (int)std::min<long long>(Tile::BM, J.m - m0);
    const int nlim = std::min(Tile::BN, J.n - n0);
    const int wlim = std::min(Tile::BN, J.ldc - n0);
    typename Tile::Acc acc;
    Tile::zero(acc);
    Tile::run((const char*)J.A + m0 * J.lda * ES, J.lda, (const char*)J.B + n0 * J.lda * ES,
              J.lda, J.kp, mlim, nlim, lds, acc);
    float* Cm = J.C + m0 * J.ldc + n0;
#pragma unroll
    for (int jj = 0; jj < Tile::TN; jj++) {
      const int n = Tile::col(jj);
#pragma unroll
      for (int i = 0; i < Tile::TM; i++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int m = Tile::row(i, r);
          if (m < mlim && n < wlim) {
            const float v = acc[i][jj][r];
            Cm[(long long)m * J.ldc + n] = v * v;
          }
        }
    }
  }
}

template <class Tile>
__global__ __launch_bounds__(Tile::NT, 1) void k_theta_lp_grouped(
    const LpThetaJob* __restrict__ jobs, int njobs, long long total) {
  __shared__ __attribute__((aligned(16))) char lds[Tile::LDS_BYTES];
  constexpr int ESA = Tile::A_PRE ? 2 : 4;
  for (long long t = blockIdx.x; t < total; t += gridDim.x) {
    int j = 0;
    while (j + 1 < njobs && jobs[j + 1].tile0 <= t) j++;
    const LpThetaJob J = jobs[j];
    const int nstrip = (J.ng + Tile::BN - 1) / Tile::BN;
    const int ntile = (J.rk + Tile::BM - 1) / Tile::BM;
    const long long local = t - J.tile0;
    if (local >= (long long)nstrip * ntile) continue;
    int s, vt;
    xcd_map((int)local, ntile, nstrip, &s, &vt);
    const long long g0 = (long long)s * Tile::BN;
    const int v0 = vt * Tile::BM;
    const int mlim = std::min(Tile::BM, J.rk - v0);
    const int nlim = (int)std::min<long long>(Tile::BN, J.ng - g0);
    typename Tile::Acc acc;
    Tile::zero(acc);
    Tile::run((const char*)J.Wf + (long long)v0 * J.ldwr * ESA, J.ldwr, J.Zt + g0 * J.ldk, J.ldk,
              J.ldk, mlim, nlim, lds, acc);
    double* Tt = J.theta + (long long)v0 * J.ldt + g0;
#pragma unroll
    for (int i = 0; i < Tile::TM; i++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = Tile::row(i, r);
#pragma unroll
        for (int jj = 0; jj < Tile::TN; jj++) {
          const int n = Tile::col(jj);
          if (m < mlim && n < nlim) Tt[(long long)m * J.ldt + n] = (double)acc[i][jj][r];
        }
      }
  }
}

// Batched helpers (blockIdx.z = instance).
__global__ __launch_bounds__(256) void k_cct_lp_batched(const CholJob* __restrict__ jobs) {
  const CholJob J = jobs[blockIdx.z];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)J.n * J.n) return;
  const int mu = (int)(t % J.n), nu = (int)(t / J.n);
  J.P[t] = (double)J.Pz[(long long)mu * J.ldz + nu];
}

__global__ __launch_bounds__(256) void k_identity_batched(const CholJob* __restrict__ jobs) {
  const CholJob J = jobs[blockIdx.z];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)J.n * J.n) return;
  J.Gt[t] = (t % J.n == t / J.n) ? 1.0 : 0.0;
}

template <bool SPLIT>
__global__ __launch_bounds__(256) void k_pack_w_batched(const CholJob* __restrict__ jobs) {
  const CholJob J = jobs[blockIdx.z];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)J.n * J.ldz) return;
  const int k = (int)(t % J.ldz), nu = (int)(t / J.ldz);
  put_operand<SPLIT>(J.Wf, J.ldz, nu, k, k < J.n ? (float)J.W[(long long)nu * J.n + k] : 0.f);
}

__global__ void k_clear_info_batched(const CholJob* __restrict__ jobs, int nb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) *jobs[i].info = 0;
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

// Tile variants.  FP32 mode: FP32 operands.  BF16X3 mode: Z C^T takes both
// operands pre-split, Theta takes W pre-split and splits Zt while staging.
typedef LpTile<256, 128, 4, 2, false> LpF32_0;              // 8 waves of 64 x 64
typedef LpTile<128, 128, 2, 2, false> LpF32_1;              // 4 waves of 64 x 64
typedef LpTile<256, 256, 4, 2, false> LpF32_2;              // 8 waves of 64 x 128
typedef LpTile<256, 128, 4, 2, true, true, true> LpZB16_0;  // 8 waves of 64 x 64
typedef LpTile<128, 256, 2, 4, true, true, true> LpZB16_1;
typedef LpTile<128, 128, 2, 2, true, true, true> LpZB16_2;  // 4 waves
typedef LpTile<256, 128, 4, 2, true, true, false> LpTB16_0;
typedef LpTile<128, 256, 2, 4, true, true, false> LpTB16_1;
typedef LpTile<128, 128, 2, 2, true, true, false> LpTB16_2;
typedef LpDmaTile<256, 256, 4, 2> LpTB16D_3;  // FP32 W and Zt, 8 waves of 64 x 128
typedef LpDmaTile<256, 256, 4, 2, 2, true> LpTB16D_4;  // FP32 W, pre-split Zt (160 KB LDS)
typedef LpDmaTile<256, 128, 4, 2> LpZB16D_3;  // FP32 psi rows, 8 waves of 64 x 64
typedef LpDmaTile<256, 256, 4, 2> LpZB16D_4;  // 8 waves of 64 x 128 (one product only)

}  // namespace

int lp_variants(int mode) { return 3; }

int lp_operand_bytes(int mode) { return mode == 2 ? 6 : 4; }

int launch_pack_psi(int mode, const double* psi, long long ld, int ng, int c0, int K, int kp,
                    void* out, hipStream_t st) {
  const dim3 grid(cdiv(ng, 64), cdiv(kp, 32));
  if (mode == 2)
    hipLaunchKernelGGL(k_pack_rows<true>, grid, dim3(256), 0, st, psi, ld, ng, c0, K, kp, out);
  else
    hipLaunchKernelGGL(k_pack_rows<false>, grid, dim3(256), 0, st, psi, ld, ng, c0, K, kp, out);
  ISDF_HIP_LAUNCHED("k_pack_rows");
  return 0;
}

int front_blocks(int ng, int rk, int kp, int* pack_blocks, int* gather_blocks) {
  *pack_blocks = cdiv(ng, 64) * cdiv(kp, 32);
  *gather_blocks = cdiv((long long)rk * kp, 256);
  return 0;
}

int launch_pack_batched(int mode, const FrontJob* jobs, int nb, long long blocks,
                        hipStream_t st) {
  if (nb > kMaxFrontJobs) return set_error(1, "batched operand packing: too many instances");
  if (blocks > 0x7fffffffLL) return set_error(1, "batched operand packing: grid too large");
  // Two workgroups per CU (256 CUs): enough loads in flight for the bandwidth the
  // packing needs, without taking every workgroup slot from the K-means beside it.
  const unsigned grid = (unsigned)std::min<long long>(blocks, 512);
  if (grid == 0) return 0;
  if (mode == 2)
    hipLaunchKernelGGL(k_pack_rows_batched<true>, dim3(grid), dim3(256), 0, st, jobs, nb,
                       (int)blocks);
  else
    hipLaunchKernelGGL(k_pack_rows_batched<false>, dim3(grid), dim3(256), 0, st, jobs, nb,
                       (int)blocks);
  ISDF_HIP_LAUNCHED("k_pack_rows_batched");
  return 0;
}

int launch_gather_batched(int mode, const FrontJob* jobs, int nb, long long blocks,
                          hipStream_t st) {
  if (nb > kMaxFrontJobs) return set_error(1, "batched operand packing: too many instances");
  if (blocks > 0x7fffffffLL) return set_error(1, "batched operand packing: grid too large");
  if (mode == 2)
    hipLaunchKernelGGL(k_gather_rows_batched<true>, dim3((unsigned)blocks), dim3(256), 0, st,
                       jobs, nb);
  else
    hipLaunchKernelGGL(k_gather_rows_batched<false>, dim3((unsigned)blocks), dim3(256), 0, st,
                       jobs, nb);
  ISDF_HIP_LAUNCHED("k_gather_rows_batched");
  return 0;
}

int launch_gather_rows(int mode, const double* psi, long long ld, const int* piv, int rk, int c0,
                       int K, int kp, void* out, hipStream_t st) {
  const dim3 grid(cdiv((long long)rk * kp, 256));
  if (mode == 2)
    hipLaunchKernelGGL(k_gather_rows<true>, grid, dim3(256), 0, st, psi, ld, piv, rk, c0, K, kp,
                       out);
  else
    hipLaunchKernelGGL(k_gather_rows<false>, grid, dim3(256), 0, st, psi, ld, piv, rk, c0, K, kp,
                       out);
  ISDF_HIP_LAUNCHED("k_gather_rows");
  return 0;
}

int lp_zct_variant(int mode, int variant, bool dual) {
  if (variant < 0) variant = mode == 1 ? 2 : 4;  // measured best on the 96^3 workloads
  // Two accumulators: 256 x 256 (FP32 and direct-to-LDS BF16X3) and the 8-wave
  // pre-split BF16X3 tiles spill; use the smaller tiles.
  if (dual && mode == 1 && variant == 2) variant = 0;
  if (dual && mode == 2 && variant < 2) variant = 2;
  if (dual && mode == 2 && variant == 4) variant = 3;
  return variant;
}

int lp_zct_op_mode(int mode, int variant, bool dual) {
  return mode == 2 && lp_zct_variant(mode, variant, dual) < 3 ? 2 : 1;
}

int launch_zct_lp(int mode, int ng, int rk, int ldz, const void* A1, const void* A2, int kp1,
                  int kp2, const void* B1, const void* B2, bool dual, float* Zt, int variant,
                  hipStream_t st, int reserve_cus, bool zsplit) {
  variant = lp_zct_variant(mode, variant, dual);
  const bool pre = mode == 2 && variant < 3;  // pre-split operands (else FP32 rows)
  const long long lda1 = pre ? 3LL * kp1 : kp1, lda2 = pre ? 3LL * kp2 : kp2;
#define ISDF_ZLP4(TL, DU, ZS, PE)                                                              \
  hipLaunchKernelGGL((k_zct_lp<TL, DU, ZS, PE>), dim3(nb), dim3(TL::NT), 0, st, ng, rk, ldz,   \
                     A1, A2, kp1, kp2, lda1, lda2, B1, B2, Zt)
#define ISDF_ZLPX(TL, ZS)                                                                      \
  do {                                                                                         \
    const long long tiles = (long long)cdiv(ng, TL::BM) * cdiv(ldz, TL::BN);                   \
    const unsigned nb = grid_leaving_cus(tiles, reserve_cus);                                  \
    if (dual && nb < tiles) ISDF_ZLP4(TL, true, ZS, true);                                     \
    else if (dual) ISDF_ZLP4(TL, true, ZS, false);                                             \
    else if (nb < tiles) ISDF_ZLP4(TL, false, ZS, true);                                       \
    else ISDF_ZLP4(TL, false, ZS, false);                                                      \
  } while (0)
#define ISDF_ZLP(TL) ISDF_ZLPX(TL, false)
#define ISDF_ZLPS(TL) ISDF_ZLPX(TL, true)
  if (zsplit && (mode != 2 || variant < 3))
    return set_error(1, "pre-split Zt needs a BF16X3 direct-to-LDS Z C^T tile");
  if (mode == 1) {
    if (variant == 1) ISDF_ZLP(LpF32_1);
    else if (variant == 2) ISDF_ZLP(LpF32_2);
    else ISDF_ZLP(LpF32_0);
  } else {
    if (variant == 1) ISDF_ZLP(LpZB16_1);
    else if (variant == 2) ISDF_ZLP(LpZB16_2);
    else if (variant == 3 && zsplit) ISDF_ZLPS(LpZB16D_3);
    else if (variant == 3) ISDF_ZLP(LpZB16D_3);
    else if (variant == 4 && zsplit) ISDF_ZLPS(LpZB16D_4);
    else if (variant == 4) ISDF_ZLP(LpZB16D_4);
    else ISDF_ZLP(LpZB16_0);
  }
#undef ISDF_ZLP
#undef ISDF_ZLPS
#undef ISDF_ZLPX
#undef ISDF_ZLP4
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

int launch_inverse(int wmode, int rk, const double* L, const double* Dinv128, double* Y,
                   double* X, double* W, int ldw, void* Wf, hipStream_t st) {
  // X = L^{-T} by recursive doubling on L^{-1} (W's buffer holds the GEMM
  // temporaries until W is written), then W = X X^T.  Before this, a
  // triangular solve of the identity (rk/16 workgroups walking the block
  // columns in sequence) took 400 us at rk 1024 on the mixed forms' critical
  // path; the doubling is log2(rk/128) levels of two batched GEMMs.
  if (int rc = launch_linv_transposed(L, rk, Dinv128, Y, W, X, st)) return rc;
  if (int rc = launch_xxt(X, rk, W, st)) return rc;
  const dim3 grid(cdiv((long long)rk * ldw, 256));
  if (wmode == 2)
    hipLaunchKernelGGL(k_pack_w<true>, grid, dim3(256), 0, st, W, rk, ldw, Wf);
  else
    hipLaunchKernelGGL(k_pack_w<false>, grid, dim3(256), 0, st, W, rk, ldw, Wf);
  ISDF_HIP_LAUNCHED("k_pack_w");
  return 0;
}

int lp_theta_variant(int mode, int variant) {
  return variant >= 0 ? variant : mode == 1 ? 2 : 4;  // measured best on the 96^3 workloads
}

int lp_theta_w_mode(int mode, int variant) {
  return mode == 2 && lp_theta_variant(mode, variant) < 3 ? 2 : 1;
}

bool lp_theta_zt_split(int mode, int variant) {
  return mode == 2 && lp_theta_variant(mode, variant) == 4;
}

int launch_theta_lp(int mode, int ng, int rk, int ldk, const void* Wf, const float* Zt,
                    double* theta, long long ldt, bool add, int variant, hipStream_t st) {
  const long long ldwr = lp_theta_w_mode(mode, variant) == 2 ? 3LL * ldk : ldk;
  variant = lp_theta_variant(mode, variant);
#define ISDF_TLP(TL)                                                                            \
  do {                                                                                          \
    const unsigned nb = cdiv(ng, TL::BN) * cdiv(rk, TL::BM);                                    \
    if (add)                                                                                    \
      hipLaunchKernelGGL((k_theta_lp<TL, true>), dim3(nb), dim3(TL::NT), 0, st, ng, rk, ldk, Wf, \
                         ldwr, Zt, theta, ldt);                                                 \
    else                                                                                        \
      hipLaunchKernelGGL((k_theta_lp<TL, false>), dim3(nb), dim3(TL::NT), 0, st, ng, rk, ldk,   \
                         Wf, ldwr, Zt, theta, ldt);                                             \
  } while (0)
  if (mode == 1) {
    if (variant == 1) ISDF_TLP(LpF32_1);
    else if (variant == 2) ISDF_TLP(LpF32_2);
    else ISDF_TLP(LpF32_0);
  } else {
    if (variant == 1) ISDF_TLP(LpTB16_1);
    else if (variant == 2) ISDF_TLP(LpTB16_2);
    else if (variant == 3) ISDF_TLP(LpTB16D_3);
    else if (variant == 4) ISDF_TLP(LpTB16D_4);
    else ISDF_TLP(LpTB16_0);
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

long long lp_zct_job_tiles(int mode, int variant, int m, int ldc) {
  if (variant < 0) variant = mode == 1 ? 2 : 0;
  int bm, bn;
  if (mode == 1) {
    bm = variant == 1 ? LpF32_1::BM : variant == 2 ? LpF32_2::BM : LpF32_0::BM;
    bn = variant == 1 ? LpF32_1::BN : variant == 2 ? LpF32_2::BN : LpF32_0::BN;
  } else {
    bm = variant == 1 ? LpZB16_1::BM : variant == 2 ? LpZB16_2::BM : LpZB16_0::BM;
    bn = variant == 1 ? LpZB16_1::BN : variant == 2 ? LpZB16_2::BN : LpZB16_0::BN;
  }
  return (long long)cdiv(m, bm) * cdiv(ldc, bn);
}

long long lp_theta_job_tiles(int mode, int variant, int ng, int rk) {
  if (variant < 0) variant = mode == 1 ? 2 : 0;
  int bm, bn;
  if (mode == 1) {
    bm = variant == 1 ? LpF32_1::BM : variant == 2 ? LpF32_2::BM : LpF32_0::BM;
    bn = variant == 1 ? LpF32_1::BN : variant == 2 ? LpF32_2::BN : LpF32_0::BN;
  } else {
    bm = variant == 1 ? LpTB16_1::BM : variant == 2 ? LpTB16_2::BM : LpTB16_0::BM;
    bn = variant == 1 ? LpTB16_1::BN : variant == 2 ? LpTB16_2::BN : LpTB16_0::BN;
  }
  return (long long)cdiv(ng, bn) * cdiv(rk, bm);
}

int launch_zct_lp_grouped(int mode, int variant, const LpGemmJob* jobs, int njobs, long long total,
                          int grid, hipStream_t st) {
  if (variant < 0) variant = mode == 1 ? 2 : 0;
#define ISDF_ZG(TL)                                                                             \
  hipLaunchKernelGGL((k_zct_lp_grouped<TL>), dim3(grid), dim3(TL::NT), 0, st, jobs, njobs, total)
  if (mode == 1) {
    if (variant == 1) ISDF_ZG(LpF32_1);
    else if (variant == 2) ISDF_ZG(LpF32_2);
    else ISDF_ZG(LpF32_0);
  } else {
    if (variant == 1) ISDF_ZG(LpZB16_1);
    else if (variant == 2) ISDF_ZG(LpZB16_2);
    else ISDF_ZG(LpZB16_0);
  }
#undef ISDF_ZG
  ISDF_HIP_LAUNCHED("k_zct_lp_grouped");
  return 0;
}

int launch_theta_lp_grouped(int mode, int variant, const LpThetaJob* jobs, int njobs,
                            long long total, int grid, hipStream_t st) {
  if (variant < 0) variant = mode == 1 ? 2 : 0;
#define ISDF_TG(TL)                                                                             \
  hipLaunchKernelGGL((k_theta_lp_grouped<TL>), dim3(grid), dim3(TL::NT), 0, st, jobs, njobs,    \
                     total)
  if (mode == 1) {
    if (variant == 1) ISDF_TG(LpF32_1);
    else if (variant == 2) ISDF_TG(LpF32_2);
    else ISDF_TG(LpF32_0);
  } else {
    if (variant == 1) ISDF_TG(LpTB16_1);
    else if (variant == 2) ISDF_TG(LpTB16_2);
    else ISDF_TG(LpTB16_0);
  }
#undef ISDF_TG
  ISDF_HIP_LAUNCHED("k_theta_lp_grouped");
  return 0;
}

int launch_cct_lp_batched(const CholJob* jobs, int nb, int max_n, hipStream_t st) {
  hipLaunchKernelGGL(k_cct_lp_batched, dim3(cdiv((long long)max_n * max_n, 256), 1, nb),
                     dim3(256), 0, st, jobs);
  ISDF_HIP_LAUNCHED("k_cct_lp_batched");
  return launch_trace_eps_batched(jobs, nb, max_n, st);
}

int launch_inverse_batched(int mode, const CholJob* jobs, int nb, int max_n, int max_ldz,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_identity_batched, dim3(cdiv((long long)max_n * max_n, 256), 1, nb),
                     dim3(256), 0, st, jobs);
  ISDF_HIP_LAUNCHED("k_identity_batched");
  if (int rc = launch_potrs_identity_batched(jobs, nb, max_n, st)) return rc;
  if (int rc = launch_xxt_batched(jobs, nb, max_n, st)) return rc;
  const dim3 grid(cdiv((long long)max_n * max_ldz, 256), 1, nb);
  if (mode == 2)
    hipLaunchKernelGGL(k_pack_w_batched<true>, grid, dim3(256), 0, st, jobs);
  else
    hipLaunchKernelGGL(k_pack_w_batched<false>, grid, dim3(256), 0, st, jobs);
  ISDF_HIP_LAUNCHED("k_pack_w_batched");
  return 0;
}

int launch_clear_info_batched(const CholJob* jobs, int nb, hipStream_t st) {
  hipLaunchKernelGGL(k_clear_info_batched, dim3(cdiv(nb, 64)), dim3(64), 0, st, jobs, nb);
  ISDF_HIP_LAUNCHED("k_clear_info_batched");
  return 0;
}

}  // namespace isdf_hip
