// C ABI entry points of libisdf_hip: the single-GPU ISDF solve.
//
// Stage order follows the reference CPU ISDF (src/common/isdf.cpp:8-127):
// rank rule, piv zeroing, K-means weights and KMEAN (getpoints_Kmeans_MPI.cpp),
// then the basis (getbasis_scalapack.cpp).  Everything stays on the device;
// only piv travels to the host, because the caller's piv is host memory.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/isdf_hip.h"
#include "isdf_internal.h"

namespace isdf_hip {

// qrcp.hip: the QRCP point selection.
int qrcp_num_pre(int rk);
size_t qrcp_bytes(long long ng, int rk);
int qrcp_select(const isdf_problem* p, const double* d_psi, long long ld, int ng, int rk,
                void* scratch, int* piv_dev, hipStream_t st);

namespace {
std::mutex g_mutex;
std::string g_last_error;
std::mutex g_err_mutex;

// Grow-only device workspace, one per device.
struct Workspace {
  void* ptr = nullptr;
  size_t bytes = 0;
};
std::vector<Workspace> g_ws;
// Device copies of the caller's host Psi and Theta for isdf_hip_solve_host.
std::vector<Workspace> g_stage;

int grow(std::vector<Workspace>& pool, size_t bytes, void** out) {
  int dev = 0;
  ISDF_HIP_TRY(hipGetDevice(&dev));
  if ((int)pool.size() <= dev) pool.resize(dev + 1);
  Workspace& w = pool[dev];
  if (w.bytes < bytes) {
    if (w.ptr) ISDF_HIP_TRY(hipFree(w.ptr));
    w.ptr = nullptr;
    w.bytes = 0;
    hipError_t e = hipMalloc(&w.ptr, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      w.ptr = nullptr;
      return set_error(ISDF_ENOMEM, "device workspace allocation of " + std::to_string(bytes) +
                                        " bytes failed: " + hipGetErrorString(e));
    }
    w.bytes = bytes;
  }
  *out = w.ptr;
  return 0;
}

int workspace(size_t bytes, void** out) { return grow(g_ws, bytes, out); }

int free_pool(std::vector<Workspace>& pool) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return ISDF_EHIP;
  if ((int)pool.size() > dev && pool[dev].ptr) {
    if (hipFree(pool[dev].ptr) != hipSuccess) return ISDF_EHIP;
    pool[dev].ptr = nullptr;
    pool[dev].bytes = 0;
  }
  return ISDF_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// glibc rand() with its default seed and libstdc++'s std::random_shuffle, for
// the npt < rk branch of KMEAN (src/common/utility.cpp:667-681).
struct GlibcRand {
  std::vector<int32_t> r;
  explicit GlibcRand(uint32_t seed) {
    r.resize(34);
    r[0] = (int32_t)seed;
    for (int i = 1; i < 31; i++) {
      int64_t v = (16807LL * r[i - 1]) % 2147483647LL;
      if (v < 0) v += 2147483647LL;
      r[i] = (int32_t)v;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    for (int i = 34; i < 344; i++) r.push_back((int32_t)((uint32_t)r[i - 31] + (uint32_t)r[i - 3]));
  }
  int next() {
    size_t i = r.size();
    int32_t v = (int32_t)((uint32_t)r[i - 31] + (uint32_t)r[i - 3]);
    r.push_back(v);
    return (int)(((uint32_t)v) >> 1);
  }
};

// Stage timing events: 0 start, 1 weights, 2 K-means, 3 selection, 4 gather
// (main stream), 5 Z C^T (main), 6/7/8 start / C C^T / end of the Cholesky
// branch (helper stream), 9 join (main), 10 solve done (main).
constexpr int kNumEvents = 11;
struct Events {
  bool on = false;
  hipEvent_t e[kNumEvents];
  int create(bool want) {
    on = want;
    if (!on) return 0;
    for (int i = 0; i < kNumEvents; i++) ISDF_HIP_TRY(hipEventCreate(&e[i]));
    return 0;
  }
  int rec(int i, hipStream_t st) {
    if (on) ISDF_HIP_TRY(hipEventRecord(e[i], st));
    return 0;
  }
  double ms(int a, int b) {
    float t = 0.f;
    if (on && hipEventElapsedTime(&t, e[a], e[b]) == hipSuccess) return (double)t;
    return 0.0;
  }
  ~Events() {
    if (on)
      for (int i = 0; i < kNumEvents; i++) (void)hipEventDestroy(e[i]);
  }
};

// Per-device helper stream (default priority) and fork/join events.  The
// C C^T -> Cholesky -> solve-operand branch depends only on the pivots, not
// on the Ng x rk Z C^T, so it runs on the helper stream beside that GEMM.
//
// Also pinned host staging for what a solve reads back (piv, the K-means
// iteration count, the unique-center count, info, eps): copies into pinned
// memory stay asynchronous, so the whole solve is queued before the host
// waits once at the end (a copy into pageable memory would block the host and
// leave the GPU idle until the rest of the solve is queued).
struct Aux {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, pack = nullptr;
  int* hpiv = nullptr;       // pinned, hpiv_cap ints
  size_t hpiv_cap = 0;
  long long* hscal = nullptr;  // pinned: [0] iters, [1] rku, [2] info, [3] eps bits, [4] npt
  // The no-allocation path: s is the caller's stream, no events, and hpiv /
  // hscal point at the caller's piv and a host array (no pinned staging).
  bool fixed = false;
};
std::vector<Aux> g_aux;

int aux_pinned(Aux* a, size_t ng) {
  if (a->fixed) return 0;
  if (!a->hscal)
    ISDF_HIP_TRY(hipHostMalloc((void**)&a->hscal, 8 * sizeof(long long), hipHostMallocDefault));
  if (a->hpiv_cap < ng) {
    if (a->hpiv) ISDF_HIP_TRY(hipHostFree(a->hpiv));
    a->hpiv = nullptr;
    a->hpiv_cap = 0;
    ISDF_HIP_TRY(hipHostMalloc((void**)&a->hpiv, sizeof(int) * ng, hipHostMallocDefault));
    a->hpiv_cap = ng;
  }
  return 0;
}

// Scratch of the no-allocation entry point (isdf_hip_solve_noalloc): a
// file-scope device array that every GPU the library runs on holds until the
// process ends.  In host code `scratch` only names the symbol; its address on
// the current device comes from hipGetSymbolAddress, once per device, and only
// that address is handed to kernels and copies.  Every call on one GPU shares
// the array (calls are serialised by the solve mutex and return synchronised).
constexpr size_t kScratchBytes = (size_t)256 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
void* g_scratch[kMaxDevices] = {};

int scratch_area(size_t bytes, void** out) {
  if (bytes > kScratchBytes)
    return set_error(ISDF_EINVAL, "the problem needs " + std::to_string(bytes) +
                                      " bytes of device workspace; isdf_hip_solve_noalloc has " +
                                      std::to_string(kScratchBytes) +
                                      " (use isdf_hip_solve, which allocates)");
  int dev = 0;
  ISDF_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(ISDF_EINVAL, "device index out of range");
  if (!g_scratch[dev]) {
    void* p = nullptr;
    ISDF_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(ISDF_EHIP, "hipGetSymbolAddress returned no address for scratch");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return 0;
}

// Fork / join of the helper stream; no-ops when the helper is the caller's
// stream itself (the no-allocation path runs every stage in order on it).
int fork_helper(Aux* a, hipStream_t st) {
  if (!a->fork) return 0;
  ISDF_HIP_TRY(hipEventRecord(a->fork, st));
  ISDF_HIP_TRY(hipStreamWaitEvent(a->s, a->fork, 0));
  return 0;
}
int join_helper(Aux* a, hipStream_t st) {
  if (!a->join) return 0;
  ISDF_HIP_TRY(hipEventRecord(a->join, a->s));
  ISDF_HIP_TRY(hipStreamWaitEvent(st, a->join, 0));
  return 0;
}

int aux_stream(Aux** out) {
  int dev = 0;
  ISDF_HIP_TRY(hipGetDevice(&dev));
  if ((int)g_aux.size() <= dev) g_aux.resize(dev + 1);
  Aux& a = g_aux[dev];
  if (!a.s) {
    ISDF_HIP_TRY(hipStreamCreateWithFlags(&a.s, hipStreamNonBlocking));
    ISDF_HIP_TRY(hipEventCreateWithFlags(&a.fork, hipEventDisableTiming));
    ISDF_HIP_TRY(hipEventCreateWithFlags(&a.join, hipEventDisableTiming));
    ISDF_HIP_TRY(hipEventCreateWithFlags(&a.pack, hipEventDisableTiming));
  }
  *out = &a;
  return 0;
}

int irank(const isdf_problem* p) {
  if (p->mu_points != 0) return p->mu_points;
  // IRound, include/common/utility.hpp:18-23.
  double a = std::sqrt((double)((p->nc1 + p->nv1) * (p->nc2 + p->nv2))) * 8.0;
  if (a > 0) return (a - (int)a < 0.5) ? (int)a : ((int)a + 1);
  return ((int)a - a < 0.5) ? (int)a : ((int)a - 1);
}

int validate(const isdf_problem* p, long long* ng_out) {
  if (!p) return set_error(ISDF_EINVAL, "problem is NULL");
  for (int d = 0; d < 3; d++) {
    if (p->ngrid[d] <= 0) return set_error(ISDF_EINVAL, "ngrid must be positive");
    if (!(p->length[d] > 0.0)) return set_error(ISDF_EINVAL, "length must be positive");
  }
  long long ng = (long long)p->ngrid[0] * p->ngrid[1] * p->ngrid[2];
  if (ng > 0x7fffffffLL) return set_error(ISDF_EINVAL, "Ng exceeds the int range of piv");
  if (p->nstate <= 0) return set_error(ISDF_EINVAL, "nstate must be positive");
  int s1 = p->nocc - p->nv1, e1 = p->nocc + p->nc1, s2 = p->nocc - p->nv2, e2 = p->nocc + p->nc2;
  if (s1 < 0 || s2 < 0 || e1 > p->nstate || e2 > p->nstate || e1 <= s1 || e2 <= s2)
    return set_error(ISDF_EINVAL,
                     "pair ranges [nocc-nv, nocc+nc) must be non-empty and inside [0, nstate)");
  if (p->kmeans_max_iter < 0) return set_error(ISDF_EINVAL, "kmeans_max_iter must be >= 0");
  if (p->ld_psi != 0 && p->ld_psi < ng) return set_error(ISDF_EINVAL, "ld_psi < Ng");
  {
    const int mode = p->precision & ~ISDF_PREC_REFINE;
    if (mode != ISDF_PREC_FP64 && mode != ISDF_PREC_FP32 && mode != ISDF_PREC_BF16X3)
      return set_error(ISDF_EINVAL, "unknown precision mode");
    if (mode == ISDF_PREC_FP64 && (p->precision & ISDF_PREC_REFINE))
      return set_error(ISDF_EINVAL, "ISDF_PREC_REFINE applies to the mixed forms only");
  }
  if (p->select != ISDF_SELECT_KMEANS && p->select != ISDF_SELECT_QRCP)
    return set_error(ISDF_EINVAL, "unknown point selection");
  int rk = irank(p);
  if (rk <= 0) return set_error(ISDF_EINVAL, "rk must be positive");
  if (rk > ng) return set_error(ISDF_EINVAL, "mu_points exceeds the number of grid points");
  if (rk > 8192) return set_error(ISDF_EINVAL, "rk above 8192 is not supported");
  *ng_out = ng;
  return 0;
}

// Tile variants forced through the environment (tuning runs); -1 = default.
int env_variant(const char* name) {
  const char* v = std::getenv(name);
  return v ? std::atoi(v) : -1;
}

// CUs the Ng-row Z C^T leaves free for the concurrent Cholesky branch (its
// workgroups then walk the tiles).  Without them every small Cholesky kernel
// waits for CUs held by the GEMM.  That matters only when the branch (a chain
// of rk/64 latency-bound steps, ~0.07 ms each; the mixed forms add the
// inverse) is about as long as the GEMM (2 Ng rk K flops at ~45 TFLOP/s FP64,
// ~140 nominal mixed); when the GEMM is longer, reserved CUs only slow it.
// Measured per solve: 64^3 FP64 17.65 ms with 48 CUs free, 17.43 with 32,
// 18.66 with 24 (the branch's first kernel then waits for the whole GEMM);
// 64^3 BF16X3 7.98 ms with 64, 8.04 with 48; 96^3 BF16X3 54.1 ms with 64,
// 53.0 with none; 96^3 FP64 and the dual 96^3 case are slower with any.
int reserve_cus(bool mixed, long long ng, int rk, int ktot) {
  if (const char* v = std::getenv("ISDF_ZCT_RESERVE_CUS")) return std::atoi(v);
  const double gemm_ms = 2.0 * (double)ng * rk * ktot / (mixed ? 140e9 : 45e9);
  const double chol_ms = (rk / 64.0) * 0.07 * (mixed ? 1.6 : 1.0);
  if (gemm_ms > 1.6 * chol_ms) return 0;
  return mixed ? 64 : 32;
}

// Basis construction of the mixed forms.  Helper stream: the operand rows of
// psi (queued at the start of the solve, beside K-means), then, from the
// pivots, C C^T + eps from a small low-precision contraction of the pivot rows
// -> FP64 Cholesky -> FP64 inverse W.  Main stream: the Z C^T contraction
// (Zt, g-major FP32) -> join -> Theta = Zt W into d_theta [-> FP64 residual,
// Theta += R W].
struct MixedBufs {
  void *A1, *A2, *B1, *B2, *Wf;
  float *Zt, *Pz;
  double *W, *Pe;
};

// Theta tile variant of a mixed solve.  The pre-split-Zt tile needs the
// Z C^T epilogue to write Zt split, which only the direct-to-LDS Z C^T tiles
// do, and the correction step writes an FP32 residual into Zt: otherwise the
// FP32-Zt tile (3) is used.
int mixed_theta_variant(int mode, bool refine, bool dual) {
  int tv = env_variant("ISDF_LP_THETA_TILE");
  if (lp_theta_zt_split(mode, tv) &&
      (refine || lp_zct_variant(mode, env_variant("ISDF_LP_ZCT_TILE"), dual) < 3))
    tv = 3;
  return tv;
}

int basis_mixed(const isdf_problem* p, int ng, int rk, int mode, bool refine, const MixedBufs& m,
                int kp1, int kp2, int ldz, double* P, double* Dinv, double* Dinv128, double* Ft,
                double* Gt, int* info_dev, double* eps_dev, double* d_theta,
                const isdf_debug* dbg, Events& ev, Aux* aux, hipStream_t st) {
  const bool dual = (p->nv1 != p->nv2 || p->nc1 != p->nc2);
  hipStream_t s2 = aux->s;
  if (int rc = fork_helper(aux, st)) return rc;
  const int zv = env_variant("ISDF_LP_ZCT_TILE");  // the format of the packed rows follows it
  const int tv = mixed_theta_variant(mode, refine, dual);
  const bool zsplit = lp_theta_zt_split(mode, tv);
  // Main stream first: Z C^T once the packed psi rows are in place.  It is
  // queued before the helper branch (rk/64 Cholesky steps of three launches
  // each, then the inverse): queued after it, Z C^T would start only once the
  // host had queued the whole branch.
  if (aux->pack) ISDF_HIP_TRY(hipStreamWaitEvent(st, aux->pack, 0));
  if (int rc = launch_zct_lp(mode, ng, rk, ldz, m.A1, m.A2, kp1, dual ? kp2 : kp1, m.B1, m.B2,
                             dual, m.Zt, zv, st,
                             aux->fixed ? 0 : reserve_cus(true, ng, rk, kp1 + (dual ? kp2 : 0)),
                             zsplit))
    return rc;
  if (int rc = ev.rec(5, st)) return rc;
  if (int rc = ev.rec(6, s2)) return rc;
  // C C^T: the contraction on the pivot rows alone (rk x rk x K).
  if (int rc = launch_zct_lp(mode, rk, rk, ldz, m.B1, m.B2, kp1, dual ? kp2 : kp1, m.B1, m.B2,
                             dual, m.Pz, zv, s2))
    return rc;
  if (int rc = launch_cct_lp(m.Pz, ldz, nullptr, rk, P, eps_dev, s2)) return rc;
  if (dbg && dbg->d_cct)
    ISDF_HIP_TRY(hipMemcpyAsync(dbg->d_cct, P, sizeof(double) * (size_t)rk * rk,
                                hipMemcpyDeviceToDevice, s2));
  if (refine)
    ISDF_HIP_TRY(hipMemcpyAsync(m.Pe, P, sizeof(double) * (size_t)rk * rk,
                                hipMemcpyDeviceToDevice, s2));
  if (int rc = ev.rec(7, s2)) return rc;
  ISDF_HIP_TRY(hipMemsetAsync(info_dev, 0, sizeof(int), s2));
  if (int rc = launch_potrf(P, rk, Dinv, info_dev, s2)) return rc;
  // Only the 128-block inverses: the mixed forms need W, not the blocked
  // solve's F / G operands (Ft and Gt serve as scratch below).
  if (int rc = launch_dinv128(P, rk, Dinv, Dinv128, s2)) return rc;
  if (int rc = launch_inverse(lp_theta_w_mode(mode, tv), rk, P, Dinv128, Ft, Gt, m.W, ldz, m.Wf,
                              s2))
    return rc;
  if (int rc = ev.rec(8, s2)) return rc;
  if (int rc = join_helper(aux, st)) return rc;
  if (int rc = ev.rec(9, st)) return rc;
  if (int rc = launch_theta_lp(mode, ng, rk, ldz, m.Wf, m.Zt, d_theta, ng, false, tv, st))
    return rc;
  if (refine) {
    if (int rc = launch_resid(ng, rk, d_theta, ng, m.Pe, m.Zt, ldz, st)) return rc;
    if (int rc = launch_theta_lp(mode, ng, rk, ldz, m.Wf, m.Zt, d_theta, ng, true, tv, st))
      return rc;
  }
  if (int rc = ev.rec(10, st)) return rc;
  return 0;
}

// noalloc: the isdf_hip_solve_noalloc form -- workspace carved from the static
// scratch array, every stage in order on `st` (no helper stream, no events),
// piv and the scalars copied straight to the caller's host memory at the end.
int solve_locked(const isdf_problem* p, const double* d_psi, double* d_theta, int* h_piv,
                 hipStream_t st, isdf_report* rep, const isdf_debug* dbg, bool noalloc = false,
                 bool exact_npt = false) {
  long long ng64 = 0;
  if (int rc = validate(p, &ng64)) return rc;
  if (!d_psi || !d_theta || !h_piv) return set_error(ISDF_EINVAL, "NULL buffer argument");
  const int ng = (int)ng64;
  const int rk = irank(p);
  const int nstate = p->nstate;
  const long long ld = p->ld_psi ? p->ld_psi : ng64;
  const int s1 = p->nocc - p->nv1, K1 = p->nv1 + p->nc1;
  const int s2 = p->nocc - p->nv2, K2 = p->nv2 + p->nc2;
  const bool dual = (p->nv1 != p->nv2 || p->nc1 != p->nc2);
  const int max_iter = p->kmeans_max_iter;
  const int nJ = (rk + kNB - 1) / kNB + 1;  // 64-blocks, plus one so 128-blocks pair up
  const int mode = p->precision & ~ISDF_PREC_REFINE;
  const bool mixed = mode != ISDF_PREC_FP64;
  // Format of the packed psi / pivot rows: the one the Z C^T tile reads.
  const int opm = mixed ? lp_zct_op_mode(mode, env_variant("ISDF_LP_ZCT_TILE"), dual) : 0;
  const bool refine = (p->precision & ISDF_PREC_REFINE) != 0;
  // Mixed forms: FP32 operand rows padded to multiples of 32 in k.
  auto pad32 = [](int v) { return (v + 31) / 32 * 32; };
  const int kp1 = pad32(K1), kp2 = dual ? pad32(K2) : 0, ldz = pad32(rk);

  // Workspace layout: K-means buffers | piv | psimuT | P | Dinv | Ft | Gt | scalars.
  size_t off = 0;
  const size_t off_km = off;
  off += align256(kmeans_bytes(ng, p->ngrid, rk, max_iter));
  const size_t off_piv = off;
  off += align256(sizeof(int) * (size_t)ng);
  const size_t off_psimu = off;
  off += align256(sizeof(double) * (size_t)nstate * rk);
  const size_t off_P = off;
  off += align256(sizeof(double) * (size_t)rk * rk);
  const size_t off_Dinv = off;
  off += align256(sizeof(double) * (size_t)nJ * kNB * kNB);
  const size_t off_D128 = off;
  off += align256(sizeof(double) * (size_t)((rk + 127) / 128) * 128 * 128);
  // The solve operands F and G, each followed by one block width of readable
  // padding: the scalar-operand solve tile reads whole rows of a block column
  // past its last valid column.
  const size_t off_Ft = off;
  off += align256(sizeof(double) * ((size_t)rk * rk + kSB));
  const size_t off_Gt = off;
  off += align256(sizeof(double) * ((size_t)rk * rk + kSB));
  const size_t off_sc = off;
  off += 256;
  const bool qrcp = p->select == ISDF_SELECT_QRCP;
  const size_t off_qr = off;
  if (qrcp) off += align256(qrcp_bytes(ng, rk));
  // Mixed forms: packed psi rows, gathered Psi(mu) rows, Zt, W (FP64), Wf, and
  // the copy of C C^T + eps I the correction step multiplies by.
  size_t off_A1 = 0, off_A2 = 0, off_B1 = 0, off_B2 = 0, off_Zt = 0, off_W = 0, off_Wf = 0,
         off_Pe = 0, off_Pz = 0;
  if (mixed) {
    const size_t eb = (size_t)lp_operand_bytes(mode);
    off_A1 = off;
    off += align256(eb * (size_t)ng * kp1);
    off_A2 = off;
    off += align256(eb * (size_t)ng * kp2);
    off_B1 = off;
    off += align256(eb * (size_t)rk * kp1);
    off_B2 = off;
    off += align256(eb * (size_t)rk * kp2);
    off_Zt = off;  // FP32, or three bf16 parts per element for the pre-split Theta tile
    off += align256((lp_theta_zt_split(mode, mixed_theta_variant(mode, refine, dual)) ? 6 : 4) *
                    (size_t)ng * ldz);
    off_W = off;
    off += align256(sizeof(double) * (size_t)rk * rk);
    off_Wf = off;
    off += align256(eb * (size_t)rk * ldz);
    off_Pz = off;
    off += align256(sizeof(float) * (size_t)rk * ldz);
    if (refine) {
      off_Pe = off;
      off += align256(sizeof(double) * (size_t)rk * rk);
    }
  }
  if (noalloc && qrcp)
    return set_error(ISDF_EINVAL, "isdf_hip_solve_noalloc selects points by K-means only");
  void* wsp = nullptr;
  if (int rc = noalloc ? scratch_area(off, &wsp) : workspace(off, &wsp)) return rc;
  char* base = (char*)wsp;
  KmeansBufs kb;
  kmeans_carve(&kb, base + off_km, ng, p->ngrid, rk, max_iter);
  int* piv_dev = (int*)(base + off_piv);
  double* psimuT = (double*)(base + off_psimu);
  double* P = (double*)(base + off_P);
  double* Dinv = (double*)(base + off_Dinv);
  double* Dinv128 = (double*)(base + off_D128);
  double* Ft = (double*)(base + off_Ft);
  double* Gt = (double*)(base + off_Gt);
  int* info_dev = (int*)(base + off_sc);
  double* eps_dev = (double*)(base + off_sc + 64);

  Aux* aux = nullptr;
  Aux own;
  long long own_scal[8] = {};
  if (noalloc) {
    own.s = st;
    own.fixed = true;
    own.hpiv = h_piv;
    own.hscal = own_scal;
    aux = &own;
  } else if (int rc = aux_stream(&aux)) {
    return rc;
  }
  MixedBufs mb{};
  if (mixed) {
    mb.A1 = base + off_A1;
    mb.A2 = dual ? base + off_A2 : mb.A1;
    mb.B1 = base + off_B1;
    mb.B2 = dual ? base + off_B2 : mb.B1;
    mb.Wf = base + off_Wf;
    mb.Zt = (float*)(base + off_Zt);
    mb.Pz = (float*)(base + off_Pz);
    mb.W = (double*)(base + off_W);
    mb.Pe = refine ? (double*)(base + off_Pe) : nullptr;
  }

  Events ev;
  if (int rc = ev.create(rep != nullptr && !noalloc)) return rc;
  if (int rc = ev.rec(0, st)) return rc;
  if (mixed) {
    // Operand rows of psi on the helper stream, beside weights and K-means.
    if (int rc = fork_helper(aux, st)) return rc;
    if (int rc = launch_pack_psi(opm, d_psi, ld, ng, s1, K1, kp1, mb.A1, aux->s)) return rc;
    if (dual)
      if (int rc = launch_pack_psi(opm, d_psi, ld, ng, s2, K2, kp2, mb.A2, aux->s)) return rc;
    if (aux->pack) ISDF_HIP_TRY(hipEventRecord(aux->pack, aux->s));
  }

  // piv zeroed first (src/common/isdf.cpp:174-175).
  std::memset(h_piv, 0, sizeof(int) * (size_t)ng);

  int npt = 0, iters = 0, rku = rk;
  bool degenerate = false, speculate = false;
  if (qrcp) {
    // QRCP point selection (src/common/isdf.cpp:105-109, getpoints_QRCP.cpp).
    if (int rc = ev.rec(1, st)) return rc;
    if (int rc = aux_pinned(aux, (size_t)ng)) return rc;
    if (int rc = qrcp_select(p, d_psi, ld, ng, rk, base + off_qr, piv_dev, st)) return rc;
    if (int rc = ev.rec(2, st)) return rc;
    ISDF_HIP_TRY(hipMemcpyAsync(aux->hpiv, piv_dev, sizeof(int) * (size_t)ng,
                                hipMemcpyDeviceToHost, st));
    aux->hscal[0] = 0;
    aux->hscal[1] = rk;
    npt = ng;
  } else {

  // 1. weights (getpoints_Kmeans_MPI.cpp:33-45).
  if (int rc = launch_weights(d_psi, ld, ng, s1, s1 + K1, s2, s2 + K2, kb.w, st)) return rc;
  if (dbg && dbg->d_weights)
    ISDF_HIP_TRY(hipMemcpyAsync(dbg->d_weights, kb.w, sizeof(double) * ng,
                                hipMemcpyDeviceToDevice, st));
  if (int rc = ev.rec(1, st)) return rc;

  // 2. K-means (utility.cpp:631-988).
  // npt, the number of points above the density threshold, picks the branch.
  // It is Ng whenever no weight falls below 1e-8 of the largest: the solve
  // assumes so and queues everything without a wait, reads npt back with its
  // last results, and runs again with npt read first (one wait; the
  // degenerate case is handled on the host) when the assumption was wrong.
  // All max_iter Lloyd iterations are queued; after convergence they return
  // at once.
  if (int rc = aux_pinned(aux, (size_t)ng)) return rc;
  speculate = !exact_npt;
  if (int rc = kmeans_prepare(&kb, p->ngrid, p->length, 1e-8, speculate ? nullptr : &npt, st))
    return rc;
  if (speculate) npt = ng;
  degenerate = npt < rk;
  if (degenerate && noalloc)
    return set_error(ISDF_EINVAL,
                     "fewer grid points above the density threshold than rk (npt " +
                         std::to_string(npt) + " < rk " + std::to_string(rk) +
                         "): isdf_hip_solve_noalloc does not take that branch; use isdf_hip_solve");
  if (!degenerate) {
    if (int rc = kmeans_setup_points(&kb, p->ngrid, p->length, npt, rk, st)) return rc;
    if (int rc = kmeans_iterate(&kb, rk, max_iter, 1e-5, nullptr, st)) return rc;
    if (int rc = ev.rec(2, st)) return rc;
    if (int rc = kmeans_finish(&kb, piv_dev, st, true)) return rc;
    if (!noalloc) {
      // Into pinned staging: stays asynchronous (the no-allocation form copies
      // at the end instead, since a copy into pageable memory blocks the host).
      ISDF_HIP_TRY(hipMemcpyAsync(aux->hpiv, piv_dev, sizeof(int) * (size_t)ng,
                                  hipMemcpyDeviceToHost, st));
      ISDF_HIP_TRY(hipMemcpyAsync(&aux->hscal[0], kb.scal + KS_ITERS, sizeof(int),
                                  hipMemcpyDeviceToHost, st));
      ISDF_HIP_TRY(hipMemcpyAsync(&aux->hscal[1], kb.scal + KS_RKU, sizeof(int),
                                  hipMemcpyDeviceToHost, st));
    }
  } else {
    // Fewer points than rk above the density threshold: the reference fills piv
    // from the threshold list and shuffles the tail (utility.cpp:667-681).
    if (int rc = ev.rec(2, st)) return rc;
    std::vector<uint8_t> fl((size_t)ng);
    ISDF_HIP_TRY(hipMemcpyAsync(fl.data(), kb.flags, (size_t)ng, hipMemcpyDeviceToHost, st));
    ISDF_HIP_TRY(hipStreamSynchronize(st));
    std::vector<int> index((size_t)ng, 0);
    int c = 0;
    for (int i = 0; i < ng; i++)
      if (fl[i]) index[c++] = i;
    int k0 = 0, k1 = 0;
    for (int i = 0; i < npt; i++) {
      if (i == index[k0]) {
        h_piv[k0] = i;
        k0 = std::min(k0 + 1, rk - 1);
      } else {
        h_piv[npt + k1] = i;
        k1++;
      }
    }
    GlibcRand rng(1u);
    int* first = h_piv + npt;
    long long len = (long long)ng - npt;
    for (long long i = 1; i < len; i++) {
      long long j = rng.next() % (i + 1);
      if (i != j) std::swap(first[i], first[j]);
    }
    ISDF_HIP_TRY(hipMemcpyAsync(piv_dev, h_piv, sizeof(int) * (size_t)ng, hipMemcpyHostToDevice,
                                st));
  }
  }  // K-means selection
  if (int rc = ev.rec(3, st)) return rc;
  if (dbg && dbg->d_labels && !degenerate && !qrcp)
    ISDF_HIP_TRY(hipMemcpyAsync(dbg->d_labels, kb.label, sizeof(int) * ng,
                                hipMemcpyDeviceToDevice, st));
  if (dbg && dbg->h_trace && max_iter > 0 && !degenerate && !qrcp) {
    // Debug output: the iteration count is needed now (one extra wait).
    int it = 0;
    ISDF_HIP_TRY(hipMemcpyAsync(&it, kb.scal + KS_ITERS, sizeof(int), hipMemcpyDeviceToHost, st));
    ISDF_HIP_TRY(hipStreamSynchronize(st));
    ISDF_HIP_TRY(hipMemcpyAsync(dbg->h_trace, kb.trace, sizeof(int) * std::max(it, 0),
                                hipMemcpyDeviceToHost, st));
  }

  if (mixed) {
    // Psi(mu) operand rows (main stream), then the two branches.
    if (int rc = launch_gather_rows(opm, d_psi, ld, piv_dev, rk, s1, K1, kp1, mb.B1, st))
      return rc;
    if (dual)
      if (int rc = launch_gather_rows(opm, d_psi, ld, piv_dev, rk, s2, K2, kp2, mb.B2, st))
        return rc;
    if (int rc = ev.rec(4, st)) return rc;
    if (int rc = basis_mixed(p, ng, rk, mode, refine, mb, kp1, kp2, ldz, P, Dinv, Dinv128, Ft,
                             Gt, info_dev, eps_dev, d_theta, dbg, ev, aux, st))
      return rc;
  } else {
    // 3. Psi(mu) gather over the caller's rk (getbasis_scalapack.cpp:249-269).
    if (int rc = launch_gather_psimu(d_psi, ld, nstate, piv_dev, rk, psimuT, st)) return rc;
    if (int rc = ev.rec(4, st)) return rc;
    // Helper stream: C C^T + eps (getbasis_scalapack.cpp:390-428) as the Z C^T
    // contraction of the pivot rows, P(mu, nu) = Z C^T(piv[mu], nu), then the
    // Cholesky (:455) and the blocked-solve operands.
    hipStream_t s2s = aux->s;
    if (int rc = fork_helper(aux, st)) return rc;
    // Main stream: Z C^T into d_theta (getbasis_scalapack.cpp:297-369).  It is
    // queued before the helper branch (rk/64 Cholesky steps of three launches
    // each): queued after it, Z C^T would start only once the host had queued
    // the whole branch.
    const int zv = zct_pick_variant(ng64, env_variant("ISDF_ZCT_TILE"));
    if (int rc = launch_zct(d_psi, ld, ng, psimuT, rk, s1, K1, s2, K2, dual, d_theta, ng, zv, st,
                            aux->fixed ? 0 : reserve_cus(false, ng64, rk, K1 + (dual ? K2 : 0))))
      return rc;
    if (int rc = ev.rec(5, st)) return rc;
    if (int rc = ev.rec(6, s2s)) return rc;
    // Same tile as the Ng-row Z C^T, so P equals its pivot rows.
    if (int rc = launch_zct(psimuT, rk, rk, psimuT, rk, s1, K1, s2, K2, dual, P, rk, zv, s2s))
      return rc;
    if (int rc = launch_trace_eps(P, rk, eps_dev, s2s)) return rc;
    if (dbg && dbg->d_cct)
      ISDF_HIP_TRY(hipMemcpyAsync(dbg->d_cct, P, sizeof(double) * (size_t)rk * rk,
                                  hipMemcpyDeviceToDevice, s2s));
    if (int rc = ev.rec(7, s2s)) return rc;
    ISDF_HIP_TRY(hipMemsetAsync(info_dev, 0, sizeof(int), s2s));
    if (int rc = launch_potrf(P, rk, Dinv, info_dev, s2s)) return rc;
    if (int rc = launch_make_fg(P, rk, Dinv, Dinv128, Ft, Gt, s2s)) return rc;
    if (int rc = ev.rec(8, s2s)) return rc;
    if (int rc = join_helper(aux, st)) return rc;
    if (int rc = ev.rec(9, st)) return rc;
    // Two right-side triangular solves in place (getbasis_scalapack.cpp:469-478).
    if (int rc = launch_potrs(d_theta, ng64, ng64, rk, Ft, Gt, env_variant("ISDF_POTRS_TILE"),
                              st))
      return rc;
    if (int rc = ev.rec(10, st)) return rc;
  }

  if (noalloc && !degenerate && !qrcp) {
    // Straight into the caller's host piv and a host array (pageable copies).
    ISDF_HIP_TRY(hipMemcpyAsync(h_piv, piv_dev, sizeof(int) * (size_t)ng, hipMemcpyDeviceToHost,
                                st));
    ISDF_HIP_TRY(hipMemcpyAsync(&aux->hscal[0], kb.scal + KS_ITERS, sizeof(int),
                                hipMemcpyDeviceToHost, st));
    ISDF_HIP_TRY(hipMemcpyAsync(&aux->hscal[1], kb.scal + KS_RKU, sizeof(int),
                                hipMemcpyDeviceToHost, st));
  }
  ISDF_HIP_TRY(hipMemcpyAsync(&aux->hscal[2], info_dev, sizeof(int), hipMemcpyDeviceToHost, st));
  ISDF_HIP_TRY(hipMemcpyAsync(&aux->hscal[3], eps_dev, sizeof(double), hipMemcpyDeviceToHost, st));
  if (speculate)
    ISDF_HIP_TRY(hipMemcpyAsync(&aux->hscal[4], kb.scal + KS_NPT, sizeof(int),
                                hipMemcpyDeviceToHost, st));
  ISDF_HIP_TRY(hipStreamSynchronize(st));
  // Some weights below the threshold: everything above was computed for a
  // grid-mode K-means that does not apply.  Solve again, npt first.
  if (speculate && (int)aux->hscal[4] != ng)
    return solve_locked(p, d_psi, d_theta, h_piv, st, rep, dbg, noalloc, true);
  const int info = (int)aux->hscal[2];
  double eps = 0.0;
  std::memcpy(&eps, &aux->hscal[3], sizeof(double));
  if (!degenerate) {
    if (aux->hpiv != h_piv) std::memcpy(h_piv, aux->hpiv, sizeof(int) * (size_t)ng);
    iters = (int)aux->hscal[0];
    rku = (int)aux->hscal[1];
  }

  if (rep) {
    rep->rk = rk;
    rep->rk_kmeans = degenerate ? rk : rku;
    rep->kmeans_iters = iters;
    rep->npt = npt;
    rep->potrf_info = info;
    rep->eps = eps;
    rep->ms_weights = ev.ms(0, 1);
    rep->ms_kmeans = ev.ms(1, 2);
    rep->ms_select = ev.ms(2, 3);
    rep->ms_gather = ev.ms(3, 4);
    rep->ms_zct = ev.ms(4, 5);
    rep->ms_cct = ev.ms(6, 7);
    rep->ms_potrf = ev.ms(7, 8);
    rep->ms_potrs = ev.ms(9, 10);
    rep->ms_total = ev.ms(0, 10);
  }
  if (info != 0) return set_error(ISDF_ENOTSPD, "Matrix is not positive definite.");
  return ISDF_OK;
}

// isdf_hip_solve_host: the caller's host Psi goes to a cached device copy
// (compact, ld Ng), the device solve runs as isdf_hip_solve does, and Theta
// comes back into the caller's host buffer.
int solve_host_locked(const isdf_problem* p, const double* h_psi, double* h_theta, int* h_piv,
                      hipStream_t st, isdf_report* rep, const isdf_debug* dbg) {
  long long ng = 0;
  if (int rc = validate(p, &ng)) return rc;
  if (!h_psi || !h_theta || !h_piv) return set_error(ISDF_EINVAL, "NULL buffer");
  const int rk = irank(p);
  const long long ld = p->ld_psi ? p->ld_psi : ng;
  const size_t psi_bytes = align256(sizeof(double) * (size_t)ng * p->nstate);
  const size_t theta_bytes = sizeof(double) * (size_t)ng * rk;
  void* base = nullptr;
  if (int rc = grow(g_stage, psi_bytes + theta_bytes, &base)) return rc;
  double* d_psi = (double*)base;
  double* d_theta = (double*)((char*)base + psi_bytes);
  ISDF_HIP_TRY(hipMemcpy2DAsync(d_psi, sizeof(double) * ng, h_psi, sizeof(double) * ld,
                                sizeof(double) * ng, p->nstate, hipMemcpyHostToDevice, st));
  isdf_problem q = *p;
  q.ld_psi = 0;
  if (int rc = solve_locked(&q, d_psi, d_theta, h_piv, st, rep, dbg)) return rc;
  ISDF_HIP_TRY(hipMemcpyAsync(h_theta, d_theta, theta_bytes, hipMemcpyDeviceToHost, st));
  ISDF_HIP_TRY(hipStreamSynchronize(st));
  return ISDF_OK;
}

}  // namespace

int solve_single_locked(const isdf_problem* p, const double* d_psi, double* d_theta, int* h_piv,
                        hipStream_t st, isdf_report* rep, const isdf_debug* dbg) {
  return solve_locked(p, d_psi, d_theta, h_piv, st, rep, dbg);
}
int problem_rank(const isdf_problem* p) { return irank(p); }
int validate_problem(const isdf_problem* p, long long* ng_out) { return validate(p, ng_out); }
std::mutex& solve_mutex() { return g_mutex; }
void release_batch_plan();
int release_fit_buffer();
int release_coulomb_buffer();

int set_error(int code, const std::string& msg) {
  std::lock_guard<std::mutex> lk(g_err_mutex);
  g_last_error = msg;
  return code;
}

}  // namespace isdf_hip

extern "C" {

int isdf_hip_solve(const isdf_problem* prob, const double* d_psi, double* d_theta, int* h_piv,
                   void* stream, isdf_report* rep, const isdf_debug* dbg) {
  std::lock_guard<std::mutex> lk(isdf_hip::g_mutex);
  isdf_hip::set_error(0, "");
  const int rc =
      isdf_hip::solve_locked(prob, d_psi, d_theta, h_piv, (hipStream_t)stream, rep, dbg);
  if (rc == ISDF_EHIP) {
    // A failed HIP call may leave work queued on either stream; drain both
    // before the workspace is reused.
    (void)hipStreamSynchronize((hipStream_t)stream);
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && (int)isdf_hip::g_aux.size() > dev &&
        isdf_hip::g_aux[dev].s)
      (void)hipStreamSynchronize(isdf_hip::g_aux[dev].s);
  }
  return rc;
}

int isdf_hip_solve_noalloc(const isdf_problem* prob, const double* d_psi, double* d_theta,
                           int* h_piv, void* stream, isdf_report* rep) {
  std::lock_guard<std::mutex> lk(isdf_hip::g_mutex);
  isdf_hip::set_error(0, "");
  const int rc = isdf_hip::solve_locked(prob, d_psi, d_theta, h_piv, (hipStream_t)stream, rep,
                                        nullptr, true);
  if (rc == ISDF_EHIP) (void)hipStreamSynchronize((hipStream_t)stream);
  return rc;
}

int isdf_hip_solve_host(const isdf_problem* prob, const double* h_psi, double* h_theta,
                        int* h_piv, void* stream, isdf_report* rep, const isdf_debug* dbg) {
  std::lock_guard<std::mutex> lk(isdf_hip::g_mutex);
  isdf_hip::set_error(0, "");
  const int rc =
      isdf_hip::solve_host_locked(prob, h_psi, h_theta, h_piv, (hipStream_t)stream, rep, dbg);
  if (rc == ISDF_EHIP) {
    (void)hipStreamSynchronize((hipStream_t)stream);
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && (int)isdf_hip::g_aux.size() > dev &&
        isdf_hip::g_aux[dev].s)
      (void)hipStreamSynchronize(isdf_hip::g_aux[dev].s);
  }
  return rc;
}

int isdf_hip_rank(const isdf_problem* prob) {
  if (!prob) return -1;
  return isdf_hip::irank(prob);
}

const char* isdf_hip_last_error(void) {
  static thread_local std::string copy;
  std::lock_guard<std::mutex> lk(isdf_hip::g_err_mutex);
  copy = isdf_hip::g_last_error;
  return copy.c_str();
}

const char* isdf_hip_version(void) { return "libisdf_hip 0.1 (gfx950)"; }

int isdf_hip_release_workspace(void) {
  std::lock_guard<std::mutex> lk(isdf_hip::g_mutex);
  isdf_hip::release_batch_plan();
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return ISDF_EHIP;
  if (isdf_hip::free_pool(isdf_hip::g_ws) != ISDF_OK ||
      isdf_hip::free_pool(isdf_hip::g_stage) != ISDF_OK ||
      isdf_hip::release_fit_buffer() != ISDF_OK ||
      isdf_hip::release_coulomb_buffer() != ISDF_OK)
    return ISDF_EHIP;
  if ((int)isdf_hip::g_aux.size() > dev && isdf_hip::g_aux[dev].hpiv) {
    if (hipHostFree(isdf_hip::g_aux[dev].hpiv) != hipSuccess) return ISDF_EHIP;
    isdf_hip::g_aux[dev].hpiv = nullptr;
    isdf_hip::g_aux[dev].hpiv_cap = 0;
  }
  return ISDF_OK;
}

}  // extern "C"
