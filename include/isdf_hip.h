/*
 * isdf_hip.h — C ABI of the MI355X (gfx950) ISDF hot path.
 *
 * This is the boundary a drop-in binding calls.  It replaces the GPU ISDF of
 * LibISDF:
 *
 *   include/isdf.hpp:37   void isdf::isdffunc::ISDF(double *d_psirow, Domain domain,
 *                             int nocc, int nstate, int nv1, int nc1, int nv2, int nc2,
 *                             int mu_points, double *d_thetaCol, int *piv,
 *                             int KmeansMaxIter_ISDF, bool check=false);
 *   src/common/isdf.cpp:132-209 (driver), src/cuda/isdf-gpu.cpp:9-403 (ISDF_GPU)
 *
 * with results that follow the reference CPU ISDF (src/common/isdf.cpp:8-127,
 * "scalapack" basis + "Kmeans" points) at one rank.  Plain C types only.
 *
 * Buffers (all column-major, owned by the caller):
 *   d_psi   device, Ng x nstate doubles, leading dimension ld_psi (0 -> Ng); read only
 *   d_theta device, Ng x rk doubles, leading dimension Ng; written
 *   h_piv   host, Ng ints; written (zeroed first, as src/common/isdf.cpp:174-175)
 * Ng = ngrid[0]*ngrid[1]*ngrid[2], grid index g = ix + ngrid[0]*(iy + ngrid[1]*iz).
 * rk = mu_points, or IRound(sqrt((nc1+nv1)(nc2+nv2)) * 8) when mu_points == 0
 * (src/common/isdf.cpp:160-164); isdf_hip_rank() returns it.
 *
 * The call is synchronous: it returns after all device work on `stream`
 * (a hipStream_t, NULL for the default stream) has finished.  One call at a
 * time per process (calls are serialised internally).
 */
#ifndef ISDF_HIP_H
#define ISDF_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  ISDF_ENOTSPD corresponds to the reference's
 * ErrorHandling("Matrix is not positive definite.")
 * (src/cpu/getbasis_scalapack.cpp:458-461, src/cuda/cusolver.cpp:136-140). */
enum {
  ISDF_OK = 0,
  ISDF_EINVAL = 1,
  ISDF_ENOTSPD = 2,
  ISDF_EHIP = 3,
  ISDF_ENOMEM = 4
};

/* Arithmetic of the Z C^T / C C^T contractions and of the Theta solve.
 * FP64 is the drop-in path and follows the reference.  The mixed forms have
 * no counterpart in the reference: K-means, the Cholesky of C C^T + eps I and
 * its inverse W stay FP64; Z C^T and Theta = Z C^T W run on low-precision MFMA
 * with FP32 accumulation.
 *   ISDF_PREC_FP32    FP32 MFMA (v_mfma_f32_16x16x4f32)
 *   ISDF_PREC_BF16X3  every FP32 operand split exactly into three bf16 parts,
 *                     six partial products on v_mfma_f32_16x16x32_bf16
 * OR ISDF_PREC_REFINE into a mixed mode for one FP64 residual-correction step
 * Theta += (Z C^T - Theta (C C^T + eps I)) W; its residual is an FP64 GEMM of
 * the size of the FP64 solve (2 Ng rk^2 flops). */
enum {
  ISDF_PREC_FP64 = 0,
  ISDF_PREC_FP32 = 1,
  ISDF_PREC_BF16X3 = 2,
  ISDF_PREC_REFINE = 16
};

typedef struct isdf_problem {
  int ngrid[3];          /* Domain::numGrid (include/common/domain.hpp:12) */
  double length[3];      /* Domain::length  (include/common/domain.hpp:10) */
  int nocc, nstate;      /* pair set S1 = [nocc-nv1, nocc+nc1), S2 = [nocc-nv2, nocc+nc2) */
  int nv1, nc1, nv2, nc2;
  int mu_points;         /* 0 -> IRound rule */
  int kmeans_max_iter;   /* KmeansMaxIter_ISDF */
  int precision;         /* ISDF_PREC_* */
  int select;            /* ISDF_SELECT_*: interpolation-point selection (single solve) */
  long long ld_psi;      /* 0 -> Ng */
} isdf_problem;

/* Interpolation-point selection, the reference's s1 argument
 * (src/common/isdf.cpp:62-74, :100-113):
 *   ISDF_SELECT_KMEANS  density-weighted K-means ("Kmeans"), the GPU ISDF's method
 *   ISDF_SELECT_QRCP    pivoted QR of the Gaussian-sketched pair products ("QRCP",
 *                       getpoints_QRCP.cpp); the first rk pivots, with the
 *                       reference's srand48(41) sketch; h_piv[rk..] hold the
 *                       remaining kept grid columns in the order the pivoting left
 *                       them, then zeros (the reference continues the
 *                       factorization past rk).  Single solves only. */
enum { ISDF_SELECT_KMEANS = 0, ISDF_SELECT_QRCP = 1 };

typedef struct isdf_report {
  int rk;            /* columns of theta */
  int rk_kmeans;     /* distinct K-means centers (src/common/utility.cpp:970-972) */
  int kmeans_iters;  /* Lloyd iterations run */
  int npt;           /* grid points with weight > 1e-8 * max weight */
  int potrf_info;    /* 0, or the LAPACK-style index of the failing pivot */
  int reserved;
  double eps;        /* 1e-10 * trace(C C^T) / rk added to the diagonal */
  /* Stage times in milliseconds (HIP events on the call's stream).  ms_potrf
   * includes the preparation of the blocked solve operands (mixed forms: and
   * the FP64 inverse W); ms_potrs is the in-place triangular-solve kernel alone
   * (mixed forms: the Theta = Z C^T W product and the optional correction). */
  double ms_weights, ms_kmeans, ms_select, ms_gather, ms_zct, ms_cct, ms_potrf, ms_potrs,
      ms_total;
} isdf_report;

/* Optional intermediate outputs for validation; any pointer may be NULL. */
typedef struct isdf_debug {
  int* h_trace;        /* host int[kmeans_max_iter]: changed labels per iteration */
  double* d_weights;   /* device double[Ng]: K-means weights */
  int* d_labels;       /* device int[Ng]: final labels */
  double* d_cct;       /* device double[rk*rk]: C C^T + eps I before the Cholesky */
} isdf_debug;

/* Full ISDF: weights -> K-means -> piv -> Psi(mu) -> Z C^T -> C C^T + eps ->
 * potrf -> Theta = Z C^T (C C^T)^{-1} written into d_theta.
 * rep and dbg may be NULL.  Returns an ISDF_* status; on failure
 * isdf_hip_last_error() describes it. */
int isdf_hip_solve(const isdf_problem* prob, const double* d_psi, double* d_theta, int* h_piv,
                   void* stream, isdf_report* rep, const isdf_debug* dbg);

/* Full ISDF on host buffers: the CPU interface of LibISDF backed by the GPU.
 * Replaces, at one rank, the non-GPU build's
 *   include/isdf.hpp:39   void isdf::isdffunc::ISDF(double *psirow, Domain domain, ...,
 *                             double *thetaCol, int *piv, int KmeansMaxIter_ISDF,
 *                             int scalblocksize=64, std::string s="scalapack",
 *                             std::string s1="Kmeans", bool check=false);
 *   src/common/isdf.cpp:8-127 (its "scalapack" + "Kmeans" path)
 * h_psi: host Ng x nstate doubles (ld_psi), read only; h_theta: host Ng x rk
 * doubles (ld Ng), written; h_piv as in isdf_hip_solve.  The library copies
 * Psi into a device buffer it keeps (grown as needed, freed by
 * isdf_hip_release_workspace), runs the device solve of isdf_hip_solve with
 * the same kernels (so Theta and piv equal that call's bit for bit), and
 * copies Theta back.  Stage times in rep cover the device solve; the copies
 * are outside them.  Synchronous. */
int isdf_hip_solve_host(const isdf_problem* prob, const double* h_psi, double* h_theta,
                        int* h_piv, void* stream, isdf_report* rep, const isdf_debug* dbg);

/* isdf_hip_solve without any device or host allocation.  No counterpart in
 * LibISDF: it is for callers that cannot allocate (code running between
 * allocation-free phases, tight serving loops, harnesses that forbid
 * hipMalloc).  It runs the kernels of isdf_hip_solve, so Theta and piv equal
 * that call's bit for bit, with these differences:
 *   - the device workspace is a static array of 256 MiB that the library holds
 *     on every GPU it runs on until the process ends; a problem whose workspace
 *     exceeds it fails with ISDF_EINVAL (FP64 fits up to about Ng = 96^3 with
 *     rk = 2048; the mixed forms need Ng x rk x 4-6 bytes more, so they fit
 *     only small problems);
 *   - every stage runs in order on `stream` (no helper stream, no events);
 *   - piv is copied straight into h_piv at the end;
 *   - K-means point selection only (prob->select = ISDF_SELECT_KMEANS), and the
 *     npt < rk branch (fewer weighted grid points than rk) fails with
 *     ISDF_EINVAL;
 *   - rep (may be NULL) gets rk, rk_kmeans, kmeans_iters, npt, potrf_info and
 *     eps; the stage times are 0.
 * All calls on one GPU share the array and it keeps the last call's contents;
 * calls are serialised inside the library and return synchronised, so calls
 * that overlapped on one GPU (from another copy of the library in the same
 * process, for instance) would race.  Synchronous. */
int isdf_hip_solve_noalloc(const isdf_problem* prob, const double* d_psi, double* d_theta,
                           int* h_piv, void* stream, isdf_report* rep);

/* rk the solve will use for this problem (the reference's rank rule). */
int isdf_hip_rank(const isdf_problem* prob);

/* Message of the last failing call on any thread ("" if none). */
const char* isdf_hip_last_error(void);

/* Result of the check=true validator. */
typedef struct isdf_check {
  double fit_error;      /* ||Z - Theta C||_F / ||Z||_F   (src/common/checkcode.cpp:62-166) */
  double z_norm;         /* ||Z||_F */
  double coulomb_error;  /* ||ijVkl - C^T M C||_F / ||ijVkl||_F (checkcode.cpp:168-312); -1 if skipped */
  double coulomb_norm;   /* ||ijVkl||_F; -1 if skipped */
} isdf_check;

/* check=true validator on host copies of psi (Ng x nstate, ld_psi), theta
 * (Ng x rk, ld Ng) and piv.  Replaces isdffunc::checkcode
 * (src/common/checkcode.cpp:54-313).  The fit error streams over pair columns;
 * the Coulomb part (do_coulomb != 0) holds (K1*K2 + rk) spectra of Ng complex
 * values and is meant for small grids. */
int isdf_hip_check(const isdf_problem* prob, const double* h_psi, const double* h_theta,
                   const int* h_piv, int do_coulomb, isdf_check* out);

/* Fit error of a result on the device: the fit-error half of check=true,
 * fit = ||Z - Theta C||_F / ||Z||_F over all Ng rows and K1*K2 pairs
 * (src/common/checkcode.cpp:62-166).  d_psi and d_theta as isdf_hip_solve
 * takes and writes them (device; Theta Ng x rk, ld Ng), h_piv host (the first
 * rk entries are used and must be grid indices).  Z is never stored: the
 * residual is formed in the epilogue of an FP64 MFMA GEMM Theta x C, so the
 * cost is 2 Ng rk K1 K2 flops and the memory O(rk) columns of C at a time.
 * Fills fit_error and z_norm; coulomb_error and coulomb_norm are set to -1.
 * Deterministic.  Synchronous.  No counterpart entry point in LibISDF, whose
 * validator runs on the host (isdf_hip_check is that one). */
int isdf_hip_fit_error(const isdf_problem* prob, const double* d_psi, const double* d_theta,
                       const int* h_piv, void* stream, isdf_check* out);

/* check=true on the device: the fit error of isdf_hip_fit_error and, with
 * do_coulomb != 0, the Coulomb error of checkcode.cpp:168-312 (same fields as
 * isdf_hip_check).  The Coulomb half takes the forward 3-D DFT of the K1*K2
 * pair densities and the rk Theta columns (direct per-axis sums), scales
 * them by sqrt(V(G)), and forms ijVkl and Theta^T V Theta as FP64 MFMA Gram
 * products over 2 Ng real rows, then ||ijVkl - C^T M C||_F.  Cost about
 * 4 (K1 K2)^2 Ng flops; device memory about 8 (2 Ng (K1 K2 + rk) + 3 (K1 K2)^2)
 * bytes (at most 65536 pairs).  No counterpart entry point in LibISDF, whose
 * validator runs on the host.  Deterministic.  Synchronous. */
int isdf_hip_check_device(const isdf_problem* prob, const double* d_psi, const double* d_theta,
                          const int* h_piv, int do_coulomb, void* stream, isdf_check* out);

/* Report of a batched solve. */
typedef struct isdf_batch_report {
  double ms_total;        /* the whole batch, HIP events on the call's stream */
  int graph_replayed;     /* 1: the call replayed the captured hipGraph of this batch shape */
  int zct_tile;           /* tile variants of the grouped Z C^T and Theta launches */
  int theta_tile;
  int reserved;
  double zct_tile_ms[4];  /* autotuner timings of each variant (first call of a shape) */
  double theta_tile_ms[4];
} isdf_batch_report;

/* Many independent ISDF instances in one call (no counterpart in LibISDF,
 * whose ISDF() takes one instance): probs[i], d_psi[i], d_theta[i], h_piv[i]
 * as in isdf_hip_solve.  Mixed forms only (ISDF_PREC_FP32 or ISDF_PREC_BF16X3,
 * the same for every instance, equal pair sets).  K-means runs per instance on
 * concurrent branches; the contractions of all instances run as persistent
 * grouped launches and the Cholesky chains as batched launches.  The first
 * call of a batch shape (problems and buffer addresses) tunes the tile
 * variants and captures the schedule in a hipGraph; later calls with the same
 * shape replay it.  Synchronous on return.  reps (nbatch entries) and brep may
 * be NULL. */
int isdf_hip_solve_batched(int nbatch, const isdf_problem* probs, const double* const* d_psi,
                           double* const* d_theta, int* const* h_piv, void* stream,
                           isdf_report* reps, isdf_batch_report* brep);

/* Library version string. */
const char* isdf_hip_version(void);

/* Releases the cached device buffers of the calling device (the solve
 * workspace, the host-interface copies, the fit-error buffer) and the batch
 * plan (buffers and captured graph). */
int isdf_hip_release_workspace(void);

#ifdef __cplusplus
}
#endif

#endif /* ISDF_HIP_H */
