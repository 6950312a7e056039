// Internal declarations shared by the HIP translation units of libisdf_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace isdf_hip {

// Records a message for isdf_hip_last_error(); returns `code`.
int set_error(int code, const std::string& msg);

#define ISDF_HIP_TRY(expr)                                                        \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::isdf_hip::set_error(3, std::string(#expr " failed: ") +            \
                                          hipGetErrorString(e_));                 \
  } while (0)

#define ISDF_HIP_LAUNCHED(name)                                                   \
  do {                                                                            \
    hipError_t e_ = hipGetLastError();                                            \
    if (e_ != hipSuccess)                                                         \
      return ::isdf_hip::set_error(3, std::string("launch of " name " failed: ") + \
                                          hipGetErrorString(e_));                 \
  } while (0)

// ---------------------------------------------------------------------------
// K-means point selection (kmeans.hip).  All FP64 arithmetic there is compiled
// with contraction off so that it reproduces the CPU reference bit for bit.
// ---------------------------------------------------------------------------

// Device buffers of one K-means run.  n = Ng points, rkmax = initial rk.
struct KmeansBufs {
  int n, rkmax, nchunk_sort, nchunk_final, chunk_sort, chunk_final, nblk_flag;
  int ngrid[3];      // grid dimensions (x fastest)
  double dx[3];      // grid spacing L/n
  int nblk_pts;      // point blocks of the assignment (bricks, or chunks of 256)
  bool grid_mode;    // every grid point is above the threshold (npt == Ng)
  double* w;         // [n]   density weights
  uint8_t* flags;    // [n]   threshold / center flags
  int* index;        // [n]   points above the threshold (ascending)
  int* gidx;         // [n]   grid index of local point i
  double *X, *Y, *Z; // [n]   point coordinates
  double* WL;        // [n]   point weights
  int* label;        // [n]
  int* sorted;       // [n]   point indices sorted stably by label
  int* bcount;       // [nblk_flag+1]
  int* hist;         // [nchunk_sort * rkmax]  chunk-major label histogram
  int* cnt_pts;      // [rkmax]
  int* base;         // [rkmax+1]
  int* map;          // [rkmax]
  double *cx, *cy, *cz;        // [rkmax] centroids
  double *sw, *sx, *sy, *sz;   // [rkmax] centroid sums
  double* pbest;     // [nchunk_final * rkmax]
  int* pidx;         // [nchunk_final * rkmax]
  int* imin;         // [rkmax]
  int* centers;      // [rkmax]
  int* trace;        // [max_iter]
  int* porder;       // [nblk_pts * 256] point of each brick slot (-1: none)
  int* changed_blk;  // [nblk_pts] labels changed per point block
  int* bbox;         // [6 * rkmax] cluster bounding boxes in grid indices
  int* cost;         // [rkmax] cost of each cluster's centroid sums (box positions + 2 members)
  int* order;        // [rkmax] clusters of the next centroid sums, costliest first
  int nsb;           // super-bricks (2 x 2 x 2 bricks) of the grid
  int* slist;        // [nsb * kSuperCap] candidate centroids of each super-brick, ascending
  double* slxyz;     // [3 * nsb * kSuperCap] their coordinates (x, y, z planes)
  int* scount;       // [nsb] their number (above kSuperCap: the list overflowed)
  int* scal;         // small device scalars, see KS_* below
};

// Capacity of a super-brick candidate list (k_cull_super).
constexpr int kSuperCap = 1024;

enum {
  KS_RK = 0,       // current number of centroids
  KS_NPT = 1,      // points above the threshold
  KS_CHANGED = 2,  // labels changed in the current iteration
  KS_ACTIVE = 3,   // 1 while the K-means loop should continue
  KS_ITERS = 4,    // iterations done
  KS_RKU = 5,      // distinct centers after unique()
  KS_MAXW = 6,     // (2 ints) bit pattern of max weight
  KS_COUNT = 8
};

// One instance of a batched K-means run (device array; the batched launches
// take blockIdx.y as the instance).  Every grid point is a K-means point.
struct KmJob {
  KmeansBufs b;
  const double* psi;
  long long ld;
  int lo, hi, s1lo, s1hi, s2lo, s2hi;  // state ranges of the weights
  int rk, max_iter;
  double maxF;                          // kmeans_tol * n
  int* piv;                             // device pivot list (n entries)
};

// Weights -> K-means -> piv for every instance, each stage one launch over all
// instances; no host synchronisation (every iteration is queued; converged
// instances return at once).  h_jobs mirrors d_jobs on the host.  d_bricks
// lists the assignment blocks of every instance, (job, brick) pairs of ints
// (nbricks pairs: every brick of job 0, then of job 1, ...); d_clusters the
// centroid-sum blocks, (job, cluster) pairs (nclusters pairs, rkmax per job).
int kmeans_batched(const KmJob* d_jobs, const KmJob* h_jobs, int nb, const int* d_bricks,
                   int nbricks, const int* d_clusters, int nclusters, hipStream_t st);
// Assignment blocks (bricks) of a grid whose every point is a K-means point.
int kmeans_grid_bricks(const int ngrid[3]);

size_t kmeans_bytes(int n, const int ngrid[3], int rkmax, int max_iter);
void kmeans_carve(KmeansBufs* b, void* base, int n, const int ngrid[3], int rkmax, int max_iter);

// Weights w(r) = (sum_{S2} psi^2)(sum_{S1} psi^2) for all grid points.
int launch_weights(const double* psi, long long ld, int ng, int s1lo, int s1hi, int s2lo,
                   int s2hi, double* w, hipStream_t st);

// Threshold, point setup, initial centroids.  Returns npt through *npt_host
// (one synchronisation); with npt_host NULL nothing is read back (npt stays in
// bcount[nblk_flag] on the device) and the caller assumes npt == Ng.
int kmeans_prepare(KmeansBufs* b, const int ngrid[3], const double length[3], double df_tol,
                   int* npt_host, hipStream_t st);
int kmeans_setup_points(KmeansBufs* b, const int ngrid[3], const double length[3], int npt,
                        int rk, hipStream_t st);
// Host-side geometry of b (grid, spacing, point blocks) for npt points.
void kmeans_geometry(KmeansBufs* b, const int ngrid[3], const double length[3], int npt);
// The Lloyd loop; returns the number of iterations through *iters_host.  With
// iters_host NULL all max_iter iterations are queued without host reads (the
// kernels stop working once converged; the count stays in scal[KS_ITERS]).
int kmeans_iterate(KmeansBufs* b, int rk, int max_iter, double kmeans_tol, int* iters_host,
                   hipStream_t st);
// Final nearest points, unique(), and the full pivot list piv_dev[n].
// nosync: no host read of the distinct-center count.
int kmeans_finish(KmeansBufs* b, int* piv_dev, hipStream_t st, bool nosync = false);

// ---------------------------------------------------------------------------
// Basis construction (basis.hip)
// ---------------------------------------------------------------------------
int launch_gather_psimu(const double* psi, long long ld, int nstate, const int* piv, int rk,
                        double* psimuT, hipStream_t st);
// `variant` < 0 picks the default tile for the shape; 0..zct_variants()-1 forces one.
int zct_variants();
// The tile variant launch_zct uses for ng rows when `requested` < 0.
int zct_pick_variant(long long ng, int requested);
// reserve_cus > 0: the default (LDS write-back) tile runs on a grid of
// (CUs - reserve_cus) workgroups that walk the tiles, leaving CUs free for work
// on another stream (the Cholesky branch beside the Ng-row product).
int launch_zct(const double* psi, long long ld, int ng, const double* psimuT, int rk, int s1,
               int K1, int s2, int K2, bool dual, double* C, long long ldc, int variant,
               hipStream_t st, int reserve_cus = 0);
// Grid for `tiles` tiles leaving `reserve` CUs free (a multiple of 8; all
// tiles when reserve <= 0).
int grid_leaving_cus(long long tiles, int reserve);
int launch_potrf(double* P, int rk, double* Dinv, int* info_dev, hipStream_t st);
int launch_make_fg(const double* L, int rk, const double* Dinv, double* Dinv128, double* Ft,
                   double* Gt, hipStream_t st);
int potrs_variants();
// T (ng x rk, ld ldt) <- T (L L^T)^{-1} in place.  unit_lower: T is the rk x rk
// identity and only the forward pass runs: T <- L^{-T} (upper triangular; the
// zero blocks are skipped).
int launch_potrs(double* T, long long ng, long long ldt, int rk, const double* Ft,
                 const double* Gt, int variant, hipStream_t st, bool unit_lower = false);

// trace of P summed in ascending order and eps = 1e-10 trace / rk added to the
// diagonal (one block).
int launch_trace_eps(double* P, int rk, double* eps_dev, hipStream_t st);

constexpr int kNB = 64;   // block size of the Cholesky
constexpr int kSB = 128;  // block-column width of the triangular solves

// One instance of a batched Cholesky / solve-operand chain (device array; the
// batched launches take blockIdx.z as the instance).  Sizes of the buffers as
// the single solve carves them for rk = n.
struct CholJob {
  double* P;        // n x n: C C^T, + eps I, then L in place (lower)
  double* Dinv;     // inverses of the 64-blocks of L
  double* Dinv128;  // inverses of the 128-blocks
  double* Ft;       // forward-solve operand (n x n)
  double* Gt;       // backward-solve operand (n x n)
  double* W;        // (L L^T)^{-1}
  void* Wf;         // operand rows of W in the mixed format (n x ldz)
  float* Pz;        // pivot-row contraction, n x ldz FP32
  int* info;        // Cholesky status
  double* eps;      // the diagonal shift
  int n, ldz;
};

// One job of the grouped low-precision Z C^T launch: C[m][c] = P1(m,c)^2 for
// c < n (c < ldc written, zero past n), P1 = A (m rows) B^T (n rows), operand
// rows of kp (padded) k values in the mode's format, row stride lda elements.
struct LpGemmJob {
  const void* A;
  const void* B;
  float* C;
  long long lda;
  int kp, m, n, ldc;
  int tile0;  // first tile of the job in the grouped tile list (a multiple of 8)
};
// One job of the grouped Theta launch: theta(g, nu) = sum_k Wf[nu][k] Zt[g][k].
struct LpThetaJob {
  const void* Wf;
  long long ldwr;  // row stride of Wf (elements of its format)
  const float* Zt;
  double* theta;
  long long ldt;
  int ldk, ng, rk;
  int tile0;
};
// Tiles of one job (before rounding the prefix to a multiple of 8).
long long lp_zct_job_tiles(int mode, int variant, int m, int ldc);
long long lp_theta_job_tiles(int mode, int variant, int ng, int rk);
// Persistent grouped launches over the jobs' tile lists; `grid` workgroups
// stride over `total` tiles.
int launch_zct_lp_grouped(int mode, int variant, const LpGemmJob* jobs, int njobs, long long total,
                          int grid, hipStream_t st);
int launch_theta_lp_grouped(int mode, int variant, const LpThetaJob* jobs, int njobs,
                            long long total, int grid, hipStream_t st);
// Batched helpers of the mixed chain: P from Pz (+ eps), W = I, Wf from W,
// info = 0.
int launch_cct_lp_batched(const CholJob* jobs, int nb, int max_n, hipStream_t st);
int launch_inverse_batched(int mode, const CholJob* jobs, int nb, int max_n, int max_ldz,
                           hipStream_t st);
int launch_clear_info_batched(const CholJob* jobs, int nb, hipStream_t st);

// Batched forms (blockIdx.z = instance): the same kernels as the single
// solve, sized for the largest instance; blocks past an instance's size exit.
int launch_trace_eps_batched(const CholJob* jobs, int nb, int max_n, hipStream_t st);
int launch_potrf_batched(const CholJob* jobs, int nb, int max_n, hipStream_t st);
int launch_make_fg_batched(const CholJob* jobs, int nb, int max_n, hipStream_t st);
// Gt <- L^{-T} per instance (Gt holds the identity; the forward pass only).
int launch_potrs_identity_batched(const CholJob* jobs, int nb, int max_n, hipStream_t st);
// W = X X^T for upper-triangular X = L^{-T}: W = (L L^T)^{-1}.
int launch_xxt(const double* X, int n, double* W, hipStream_t st);
// The 128-block inverses of the factor alone (the first half of launch_make_fg).
int launch_dinv128(const double* L, int rk, const double* Dinv, double* Dinv128, hipStream_t st);
// X = L^{-T} (n x n, column-major) by recursive doubling on L^{-1} from the
// 128-block inverses; Y and Tb are n x n scratch.
int launch_linv_transposed(const double* L, int n, const double* Dinv128, double* Y, double* Tb,
                           double* X, hipStream_t st);
int launch_xxt_batched(const CholJob* jobs, int nb, int max_n, hipStream_t st);

// ---------------------------------------------------------------------------
// Mixed-precision contractions (lowprec.hip).  mode 1 = FP32 MFMA, mode 2 =
// 3-way-split BF16 MFMA; FP32 accumulation in both.  Operands are FP32 rows
// with k contiguous, k padded to a multiple of 32 with zeros.
// ---------------------------------------------------------------------------
int lp_variants(int mode);
// Bytes per operand element: 4 (FP32 rows) or 6 (BF16X3: three pre-split bf16
// planes interleaved per 32-wide k chunk).
int lp_operand_bytes(int mode);
// Operand rows g of psi(g, c0 + k), k < kp (zero for k >= K), in the mode's format.
int launch_pack_psi(int mode, const double* psi, long long ld, int ng, int c0, int K, int kp,
                    void* out, hipStream_t st);
// Operand rows nu of psi(piv[nu], c0 + k) in the mode's format.
// One instance of the batched operand packing: psi columns c0 .. c0 + K - 1 of
// its grid rows into A (ng x kp) and of its pivot rows into B (rk x kp), in the
// format of the mode; pack_b0 / gather_b0 are the job's first blocks in the two
// batched grids (front_blocks gives each job's block counts).
constexpr int kMaxFrontJobs = 1024;
struct FrontJob {
  const double* psi;
  long long ld;
  const int* piv;
  void* A;
  void* B;
  int ng, rk, c0, K, kp;
  int pack_b0, gather_b0;
};
int front_blocks(int ng, int rk, int kp, int* pack_blocks, int* gather_blocks);
// The grid rows of every instance (A; needs psi only) and its pivot rows (B;
// needs piv), one launch each.
int launch_pack_batched(int mode, const FrontJob* jobs, int nb, long long blocks,
                        hipStream_t st);
int launch_gather_batched(int mode, const FrontJob* jobs, int nb, long long blocks,
                          hipStream_t st);
int launch_gather_rows(int mode, const double* psi, long long ld, const int* piv, int rk, int c0,
                       int K, int kp, void* out, hipStream_t st);
// Zt[g*ldz + nu] = P1(g,nu) P2(g,nu) (FP32), zero for rk <= nu < ldz.
// zsplit: Zt written pre-split (bf16 parts, row stride 3 ldz; BF16X3 variants
// 3 and 4 only), for Theta variant 4.
int launch_zct_lp(int mode, int ng, int rk, int ldz, const void* A1, const void* A2, int kp1,
                  int kp2, const void* B1, const void* B2, bool dual, float* Zt, int variant,
                  hipStream_t st, int reserve_cus = 0, bool zsplit = false);
// P(mu, nu) = Zt[piv[mu]][nu] (column-major FP64; piv NULL: row mu) and the eps shift.
int launch_cct_lp(const float* Zt, int ldz, const int* piv, int rk, double* P, double* eps_dev,
                  hipStream_t st);
// W = (L L^T)^{-1} (FP64, column-major rk x rk) from the factor L and its
// 128-block inverses (Y and X are rk x rk scratch: L^{-1} and L^{-T} on the
// way), and operand rows nu of W(k, nu), k < ldw (zero for k >= rk), in
// operand format wmode (1 FP32, 2 pre-split bf16 planes).
int launch_inverse(int wmode, int rk, const double* L, const double* Dinv128, double* Y,
                   double* X, double* W, int ldw, void* Wf, hipStream_t st);
// Z C^T tile variant after defaults and the dual-product overrides, and the
// format of the operand rows it reads (1 FP32, 2 pre-split bf16 planes; the
// mode argument of launch_pack_psi / launch_gather_rows): BF16X3 variants 3
// and 4 read FP32 rows and split them in registers.
int lp_zct_variant(int mode, int variant, bool dual);
int lp_zct_op_mode(int mode, int variant, bool dual);
// Theta tile variant (-1: the default of the mode) and the W operand format it
// reads (the wmode of launch_inverse): BF16X3 variant 3 reads FP32 W.
int lp_theta_variant(int mode, int variant);
int lp_theta_w_mode(int mode, int variant);
// Whether the Theta variant reads Zt pre-split (BF16X3 variant 4).
bool lp_theta_zt_split(int mode, int variant);
// theta(g, nu) = (or +=) sum_k Wf[nu][k] Zt[g][k], k < ldk.
int launch_theta_lp(int mode, int ng, int rk, int ldk, const void* Wf, const float* Zt,
                    double* theta, long long ldt, bool add, int variant, hipStream_t st);
// Zt[g][nu] <- Zt[g][nu] - (theta Pe)(g, nu) in FP64 (the correction step's residual).
int launch_resid(int ng, int rk, const double* theta, long long ldt, const double* Pe, float* Zt,
                 int ldz, hipStream_t st);

}  // namespace isdf_hip
