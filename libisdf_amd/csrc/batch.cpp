// Batched ISDF: many independent instances of uneven size in one call, for
// the mixed-precision forms.  LibISDF has no batched entry point; each
// instance computes what isdf_hip_solve computes for it.
//
// Schedule of one batch (captured once into two hipGraphs and replayed on
// later calls with the same shapes and buffers):
//   1. (first graph) weights -> K-means -> piv of every instance, each stage
//      one launch over the whole batch (every Lloyd iteration queued,
//      converged instances return at once; the per-iteration kernels take
//      (job, brick) / (job, cluster) block lists), with the grid rows of psi
//      of every instance packed beside it on a helper stream;
//   -- the copies of piv to pinned host memory run between the two graphs on a
//      side stream, and the host hands piv on while the second graph runs;
//   2. (second graph) the pivot rows of every instance in one launch, one
//      persistent grouped launch of the pivot-row contractions (C C^T of every
//      instance), then the batched Cholesky chain (every launch covers all
//      instances: blockIdx.z = instance) up to each instance's W; beside it, on
//      a helper stream, one persistent grouped launch of every instance's Z C^T;
//   3. one persistent grouped launch of every instance's Theta = Zt W.
// The first call for a batch shape runs the schedule eagerly and times every
// tile variant of the two grouped contractions (the autotuner); the graphs are
// captured with the fastest pair.  K-means is assumed to see every grid point
// (npt == Ng); an instance where that fails is solved again by the single
// solve, which handles any npt.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/isdf_hip.h"
#include "isdf_internal.h"

namespace isdf_hip {

// From solve.cpp.
int solve_single_locked(const isdf_problem* p, const double* d_psi, double* d_theta, int* h_piv,
                        hipStream_t st, isdf_report* rep, const isdf_debug* dbg);
int problem_rank(const isdf_problem* p);
int validate_problem(const isdf_problem* p, long long* ng_out);
std::mutex& solve_mutex();

namespace {

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
inline int pad32(int v) { return (v + 31) / 32 * 32; }

struct InstStatus {
  int npt, iters, rku, info;
  double eps;
};

// Per-instance device status gathered at the end of the schedule.
struct StatusSrc {
  const int* npt;   // bcount[nblk_flag]
  const int* scal;  // K-means scalars
  const int* info;
  const double* eps;
};

__global__ void k_batch_status(const StatusSrc* __restrict__ src, int nb,
                               InstStatus* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  InstStatus s;
  s.npt = *src[i].npt;
  s.iters = src[i].scal[KS_ITERS];
  s.rku = src[i].scal[KS_RKU];
  s.info = *src[i].info;
  s.eps = *src[i].eps;
  out[i] = s;
}

struct Inst {
  isdf_problem p;
  const double* psi;
  double* theta;
  int ng, rk, s1, K1, kp, ldz;
  long long ld;
  KmeansBufs kb;
  int* piv_dev;
  void *A, *B, *Wf;
  float *Zt, *Pz;
  double *P, *Dinv, *D128, *Ft, *Gt, *W, *eps;
  int* info;
  size_t piv_host_off;
};

struct Plan {
  std::vector<Inst> inst;
  int mode = 0, nb = 0, max_rk = 0, max_ldz = 0, grid = 0;
  int zv = -1, tv = -1;           // tile variants chosen by the tuner
  double zms[4] = {0}, tms[4] = {0};
  void* dev = nullptr;            // all device memory of the batch
  size_t dev_bytes = 0;
  int* piv_host = nullptr;        // pinned
  InstStatus* status_host = nullptr;  // pinned
  CholJob* d_chol = nullptr;
  LpGemmJob *d_pjobs = nullptr, *d_zjobs = nullptr;
  LpThetaJob* d_tjobs = nullptr;
  StatusSrc* d_src = nullptr;
  InstStatus* d_status = nullptr;
  KmJob* d_km = nullptr;
  int* d_bricks = nullptr;           // (job, brick) of every K-means assignment block
  int nbricks = 0;
  int* d_clusters = nullptr;         // (job, cluster) of every centroid-sum block
  int nclusters = 0;
  FrontJob* d_front = nullptr;       // the batched operand packing
  long long front_pack = 0, front_gather = 0;  // its two grids (blocks)
  std::vector<KmJob> h_km;
  std::vector<LpGemmJob> pjobs, zjobs;
  std::vector<LpThetaJob> tjobs;
  long long ptotal = 0, ztotal = 0, ttotal = 0;
  std::vector<hipStream_t> hs;
  std::vector<hipEvent_t> ev;  // kEvCount events (the enum below)
  // Recorded once every piv is in piv_host: the host copies piv out to the caller
  // while the GPU runs the rest.
  hipEvent_t piv_ready = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};     // the front and back parts
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  hipStream_t copy_stream = nullptr;             // piv copies beside the back part
  int device = -1;

  void release() {
    for (int p = 0; p < 2; p++) {
      if (exec[p]) (void)hipGraphExecDestroy(exec[p]);
      if (graph[p]) (void)hipGraphDestroy(graph[p]);
      exec[p] = nullptr;
      graph[p] = nullptr;
    }
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    copy_stream = nullptr;
    for (auto s : hs) (void)hipStreamDestroy(s);
    for (auto e : ev) (void)hipEventDestroy(e);
    if (piv_ready) (void)hipEventDestroy(piv_ready);
    piv_ready = nullptr;
    hs.clear();
    ev.clear();
    if (dev) (void)hipFree(dev);
    if (piv_host) (void)hipHostFree(piv_host);
    if (status_host) (void)hipHostFree(status_host);
    dev = nullptr;
    piv_host = nullptr;
    status_host = nullptr;
  }
};

Plan g_plan;
constexpr int kHelpers = 2;  // helper streams: Z C^T beside the Cholesky chain; the piv copies
// Events of the schedule.
enum { kEvFork, kEvPackFork, kEvPackDone, kEvZFork, kEvZJoin, kEvCount };

bool same_batch(const Plan& pl, int nb, const isdf_problem* probs, const double* const* psi,
                double* const* theta, int dev) {
  if (pl.nb != nb || pl.device != dev || !pl.dev) return false;
  for (int i = 0; i < nb; i++) {
    if (std::memcmp(&pl.inst[i].p, &probs[i], sizeof(isdf_problem)) != 0) return false;
    if (pl.inst[i].psi != psi[i] || pl.inst[i].theta != theta[i]) return false;
  }
  return true;
}

int build_plan(Plan& pl, int nb, const isdf_problem* probs, const double* const* psi,
               double* const* theta, int dev) {
  pl.release();
  pl.nb = nb;
  pl.device = dev;
  pl.mode = probs[0].precision & ~ISDF_PREC_REFINE;
  pl.inst.assign(nb, Inst{});
  pl.max_rk = 0;
  pl.max_ldz = 0;
  const size_t eb = (size_t)lp_operand_bytes(pl.mode);
  // Layout: per instance [K-means | piv | A | B | Zt | Pz | Wf | P | Dinv | D128 | Ft | Gt | W |
  // scalars], then the job arrays.
  size_t off = 0, piv_total = 0;
  long long bricks = 0, clusters = 0;
  std::vector<size_t> base(nb);
  for (int i = 0; i < nb; i++) {
    Inst& I = pl.inst[i];
    I.p = probs[i];
    I.psi = psi[i];
    I.theta = theta[i];
    I.ng = probs[i].ngrid[0] * probs[i].ngrid[1] * probs[i].ngrid[2];
    I.rk = problem_rank(&probs[i]);
    I.s1 = probs[i].nocc - probs[i].nv1;
    I.K1 = probs[i].nv1 + probs[i].nc1;
    I.kp = pad32(I.K1);
    I.ldz = pad32(I.rk);
    I.ld = probs[i].ld_psi ? probs[i].ld_psi : I.ng;
    pl.max_rk = std::max(pl.max_rk, I.rk);
    pl.max_ldz = std::max(pl.max_ldz, I.ldz);
    const int nJ = (I.rk + kNB - 1) / kNB + 1;
    base[i] = off;
    off += al256(kmeans_bytes(I.ng, probs[i].ngrid, I.rk, probs[i].kmeans_max_iter));
    off += al256(sizeof(int) * (size_t)I.ng);
    off += al256(eb * (size_t)I.ng * I.kp);
    off += al256(eb * (size_t)I.rk * I.kp);
    off += al256(sizeof(float) * (size_t)I.ng * I.ldz);
    off += al256(sizeof(float) * (size_t)I.rk * I.ldz);
    off += al256(eb * (size_t)I.rk * I.ldz);
    off += al256(sizeof(double) * (size_t)I.rk * I.rk);
    off += al256(sizeof(double) * (size_t)nJ * kNB * kNB);
    off += al256(sizeof(double) * (size_t)((I.rk + 127) / 128) * 128 * 128);
    off += 3 * al256(sizeof(double) * (size_t)I.rk * I.rk);
    off += 256;
    I.piv_host_off = piv_total;
    piv_total += (size_t)I.ng;
    bricks += kmeans_grid_bricks(probs[i].ngrid);
    clusters += I.rk;
  }
  if (bricks > 0x7fffffffLL || clusters > 0x7fffffffLL)
    return set_error(ISDF_EINVAL, "batch K-means exceeds the int range of blocks");
  pl.nbricks = (int)bricks;
  pl.nclusters = (int)clusters;
  const size_t off_jobs = off;
  off += al256(sizeof(CholJob) * nb) + 2 * al256(sizeof(LpGemmJob) * nb) +
         al256(sizeof(LpThetaJob) * nb) + al256(sizeof(StatusSrc) * nb) +
         al256(sizeof(InstStatus) * nb) + al256(sizeof(KmJob) * nb) +
         al256(sizeof(FrontJob) * nb) + al256(sizeof(int) * 2 * (size_t)pl.nbricks) +
         al256(sizeof(int) * 2 * (size_t)pl.nclusters);
  hipError_t e = hipMalloc(&pl.dev, off);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    pl.dev = nullptr;
    return set_error(ISDF_ENOMEM, "batch workspace allocation of " + std::to_string(off) +
                                      " bytes failed: " + hipGetErrorString(e));
  }
  pl.dev_bytes = off;
  ISDF_HIP_TRY(hipHostMalloc((void**)&pl.piv_host, sizeof(int) * std::max<size_t>(piv_total, 1),
                             hipHostMallocDefault));
  ISDF_HIP_TRY(hipHostMalloc((void**)&pl.status_host, sizeof(InstStatus) * nb,
                             hipHostMallocDefault));
  char* d = (char*)pl.dev;
  std::vector<CholJob> chol(nb);
  std::vector<StatusSrc> src(nb);
  pl.h_km.assign(nb, KmJob{});
  for (int i = 0; i < nb; i++) {
    Inst& I = pl.inst[i];
    const int nJ = (I.rk + kNB - 1) / kNB + 1;
    size_t o = base[i];
    auto take = [&](size_t bytes) {
      char* r = d + o;
      o += al256(bytes);
      return r;
    };
    char* km = take(kmeans_bytes(I.ng, I.p.ngrid, I.rk, I.p.kmeans_max_iter));
    kmeans_carve(&I.kb, km, I.ng, I.p.ngrid, I.rk, I.p.kmeans_max_iter);
    I.piv_dev = (int*)take(sizeof(int) * (size_t)I.ng);
    I.A = take(eb * (size_t)I.ng * I.kp);
    I.B = take(eb * (size_t)I.rk * I.kp);
    I.Zt = (float*)take(sizeof(float) * (size_t)I.ng * I.ldz);
    I.Pz = (float*)take(sizeof(float) * (size_t)I.rk * I.ldz);
    I.Wf = take(eb * (size_t)I.rk * I.ldz);
    I.P = (double*)take(sizeof(double) * (size_t)I.rk * I.rk);
    I.Dinv = (double*)take(sizeof(double) * (size_t)nJ * kNB * kNB);
    I.D128 = (double*)take(sizeof(double) * (size_t)((I.rk + 127) / 128) * 128 * 128);
    I.Ft = (double*)take(sizeof(double) * (size_t)I.rk * I.rk);
    I.Gt = (double*)take(sizeof(double) * (size_t)I.rk * I.rk);
    I.W = (double*)take(sizeof(double) * (size_t)I.rk * I.rk);
    char* sc = take(256);
    I.info = (int*)sc;
    I.eps = (double*)(sc + 64);
    chol[i] = CholJob{I.P, I.Dinv, I.D128, I.Ft, I.Gt, I.W, I.Wf, I.Pz, I.info, I.eps, I.rk,
                      I.ldz};
    src[i] = StatusSrc{I.kb.scal + KS_NPT, I.kb.scal, I.info, I.eps};
    kmeans_geometry(&I.kb, I.p.ngrid, I.p.length, I.ng);
    const int s2 = I.p.nocc - I.p.nv2, K2 = I.p.nv2 + I.p.nc2;
    KmJob& kj = pl.h_km[i];
    kj.b = I.kb;
    kj.psi = I.psi;
    kj.ld = I.ld;
    kj.lo = std::min(I.s1, s2);
    kj.hi = std::max(I.s1 + I.K1, s2 + K2);
    kj.s1lo = I.s1;
    kj.s1hi = I.s1 + I.K1;
    kj.s2lo = s2;
    kj.s2hi = s2 + K2;
    kj.rk = I.rk;
    kj.max_iter = I.p.kmeans_max_iter;
    kj.maxF = 1e-5 * I.ng;
    kj.piv = I.piv_dev;
  }
  // The batched K-means kernels take their job from blockIdx.y (set-up) or
  // from the block lists below (Lloyd iterations); every job carries its own
  // buffers, so the order is free.  Largest clusters (Ng / rk) first: the
  // centroid sums of those instances are the longest dependent chains, and
  // workgroups start in grid order.
  std::stable_sort(pl.h_km.begin(), pl.h_km.end(), [](const KmJob& a, const KmJob& b) {
    return (double)a.b.n / a.rk > (double)b.b.n / b.rk;
  });
  size_t o = off_jobs;
  pl.d_chol = (CholJob*)(d + o);
  o += al256(sizeof(CholJob) * nb);
  pl.d_pjobs = (LpGemmJob*)(d + o);
  o += al256(sizeof(LpGemmJob) * nb);
  pl.d_zjobs = (LpGemmJob*)(d + o);
  o += al256(sizeof(LpGemmJob) * nb);
  pl.d_tjobs = (LpThetaJob*)(d + o);
  o += al256(sizeof(LpThetaJob) * nb);
  pl.d_src = (StatusSrc*)(d + o);
  o += al256(sizeof(StatusSrc) * nb);
  pl.d_status = (InstStatus*)(d + o);
  o += al256(sizeof(InstStatus) * nb);
  pl.d_km = (KmJob*)(d + o);
  o += al256(sizeof(KmJob) * nb);
  pl.d_front = (FrontJob*)(d + o);
  o += al256(sizeof(FrontJob) * nb);
  pl.d_bricks = (int*)(d + o);
  o += al256(sizeof(int) * 2 * (size_t)pl.nbricks);
  pl.d_clusters = (int*)(d + o);
  {
    // Block lists of the batched K-means, job by job in the order of d_km.
    std::vector<int> bl, cl;
    bl.reserve(2 * (size_t)pl.nbricks);
    cl.reserve(2 * (size_t)pl.nclusters);
    for (int j = 0; j < nb; j++) {
      for (int x = 0; x < pl.h_km[j].b.nblk_pts; x++) {
        bl.push_back(j);
        bl.push_back(x);
      }
      for (int c = 0; c < pl.h_km[j].b.rkmax; c++) {
        cl.push_back(j);
        cl.push_back(c);
      }
    }
    if (bl.size() != 2 * (size_t)pl.nbricks || cl.size() != 2 * (size_t)pl.nclusters)
      return set_error(ISDF_EINVAL, "batch K-means block count mismatch");
    ISDF_HIP_TRY(hipMemcpy(pl.d_bricks, bl.data(), sizeof(int) * bl.size(),
                           hipMemcpyHostToDevice));
    ISDF_HIP_TRY(hipMemcpy(pl.d_clusters, cl.data(), sizeof(int) * cl.size(),
                           hipMemcpyHostToDevice));
  }
  std::vector<FrontJob> front(nb);
  pl.front_pack = pl.front_gather = 0;
  for (int i = 0; i < nb; i++) {
    const Inst& I = pl.inst[i];
    int pb = 0, gb = 0;
    front_blocks(I.ng, I.rk, I.kp, &pb, &gb);
    front[i] = FrontJob{I.psi, I.ld, I.piv_dev, I.A, I.B, I.ng, I.rk, I.s1, I.K1, I.kp,
                        (int)pl.front_pack, (int)pl.front_gather};
    pl.front_pack += pb;
    pl.front_gather += gb;
    if (pl.front_pack > 0x7fffffffLL || pl.front_gather > 0x7fffffffLL)
      return set_error(ISDF_EINVAL, "batch operand packing exceeds the int range");
  }
  ISDF_HIP_TRY(hipMemcpy(pl.d_front, front.data(), sizeof(FrontJob) * nb, hipMemcpyHostToDevice));
  ISDF_HIP_TRY(hipMemcpy(pl.d_km, pl.h_km.data(), sizeof(KmJob) * nb, hipMemcpyHostToDevice));
  ISDF_HIP_TRY(hipMemcpy(pl.d_chol, chol.data(), sizeof(CholJob) * nb, hipMemcpyHostToDevice));
  ISDF_HIP_TRY(hipMemcpy(pl.d_src, src.data(), sizeof(StatusSrc) * nb, hipMemcpyHostToDevice));
  int cus = 0;
  ISDF_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  pl.grid = std::max(8, cus / 8 * 8);  // one resident workgroup per CU, a multiple of 8
  const int nstreams = std::min(kHelpers, nb);
  pl.hs.resize(nstreams);
  for (auto& s : pl.hs) ISDF_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  pl.ev.resize(kEvCount);
  for (auto& ev : pl.ev) ISDF_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  ISDF_HIP_TRY(hipEventCreateWithFlags(&pl.piv_ready, hipEventDisableTiming));
  ISDF_HIP_TRY(hipStreamCreateWithFlags(&pl.copy_stream, hipStreamNonBlocking));
  return 0;
}

// Tile lists of the grouped contractions for tile variants (zv, tv); the
// pivot-row contractions always use the default tile.  Synchronous copies: not
// while capturing, and not while a launch that reads the lists may run.
int set_jobs(Plan& pl, int zv, int tv) {
  const int nb = pl.nb;
  pl.pjobs.resize(nb);
  pl.zjobs.resize(nb);
  pl.tjobs.resize(nb);
  long long pt = 0, zt = 0, tt = 0;
  const long long lmul = pl.mode == 2 ? 3 : 1;
  for (int i = 0; i < nb; i++) {
    const Inst& I = pl.inst[i];
    pl.pjobs[i] = LpGemmJob{I.B, I.B, I.Pz, lmul * I.kp, I.kp, I.rk, I.rk, I.ldz, (int)pt};
    pl.zjobs[i] = LpGemmJob{I.A, I.B, I.Zt, lmul * I.kp, I.kp, I.ng, I.rk, I.ldz, (int)zt};
    pl.tjobs[i] = LpThetaJob{I.Wf, lmul * I.ldz, I.Zt, I.theta, (long long)I.ng, I.ldz, I.ng,
                             I.rk, (int)tt};
    pt += (lp_zct_job_tiles(pl.mode, -1, I.rk, I.ldz) + 7) / 8 * 8;
    zt += (lp_zct_job_tiles(pl.mode, zv, I.ng, I.ldz) + 7) / 8 * 8;
    tt += (lp_theta_job_tiles(pl.mode, tv, I.ng, I.rk) + 7) / 8 * 8;
  }
  if (zt > 0x7fffffffLL || tt > 0x7fffffffLL)
    return set_error(ISDF_EINVAL, "batch tile list exceeds the int range");
  pl.ptotal = pt;
  pl.ztotal = zt;
  pl.ttotal = tt;
  ISDF_HIP_TRY(hipMemcpy(pl.d_pjobs, pl.pjobs.data(), sizeof(LpGemmJob) * nb,
                         hipMemcpyHostToDevice));
  ISDF_HIP_TRY(hipMemcpy(pl.d_zjobs, pl.zjobs.data(), sizeof(LpGemmJob) * nb,
                         hipMemcpyHostToDevice));
  ISDF_HIP_TRY(hipMemcpy(pl.d_tjobs, pl.tjobs.data(), sizeof(LpThetaJob) * nb,
                         hipMemcpyHostToDevice));
  return 0;
}


struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
  int start(hipStream_t st) {
    if (!a) {
      ISDF_HIP_TRY(hipEventCreate(&a));
      ISDF_HIP_TRY(hipEventCreate(&b));
    }
    ISDF_HIP_TRY(hipEventRecord(a, st));
    return 0;
  }
  int stop(hipStream_t st, double* ms) {
    ISDF_HIP_TRY(hipEventRecord(b, st));
    ISDF_HIP_TRY(hipEventSynchronize(b));
    float t = 0.f;
    ISDF_HIP_TRY(hipEventElapsedTime(&t, a, b));
    *ms = t;
    return 0;
  }
  ~Timer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

// The whole schedule on stream m.  tune: time every tile variant of the two
// grouped contractions on this data and keep the fastest (synchronises; the
// results are those of the last variant run, identical for every variant:
// the k order of each sum does not depend on the tile).  Otherwise the job
// lists must already be set for (zv, tv); this form is the one captured.
// The batch schedule in two parts, each captured as its own graph:
//  front: the grid rows of every instance (psi only) packed on a helper stream
//         beside the K-means of every instance;
//  back:  pivot rows, the contractions, the Cholesky chains, Theta, status.
// Between them the piv copies to the host run outside the graphs (issue_piv_copies),
// beside the back part, so that the host can copy piv on to the caller while the
// GPU runs the rest (a host-waitable event inside a captured graph is not available).
int enqueue_front(Plan& pl, hipStream_t m) {
  const int nb = pl.nb, nh = (int)pl.hs.size();
  hipStream_t ps = nh > 1 ? pl.hs[nh - 1] : m;
  if (nh > 1) {
    ISDF_HIP_TRY(hipEventRecord(pl.ev[kEvPackFork], m));
    ISDF_HIP_TRY(hipStreamWaitEvent(ps, pl.ev[kEvPackFork], 0));
  }
  if (int rc = launch_pack_batched(pl.mode, pl.d_front, nb, pl.front_pack, ps)) return rc;
  if (nh > 1) ISDF_HIP_TRY(hipEventRecord(pl.ev[kEvPackDone], ps));
  // K-means of every instance, each stage one launch over the batch.
  if (int rc = kmeans_batched(pl.d_km, pl.h_km.data(), nb, pl.d_bricks, pl.nbricks, pl.d_clusters,
                                pl.nclusters, m)) return rc;
  if (nh > 1) ISDF_HIP_TRY(hipStreamWaitEvent(m, pl.ev[kEvPackDone], 0));
  return 0;
}

// piv of every instance to the pinned staging, on stream s once `after` is done;
// records pl.piv_ready.
int issue_piv_copies(Plan& pl, hipStream_t s, hipEvent_t after) {
  if (after) ISDF_HIP_TRY(hipStreamWaitEvent(s, after, 0));
  for (int i = 0; i < pl.nb; i++) {
    const Inst& I = pl.inst[i];
    ISDF_HIP_TRY(hipMemcpyAsync(pl.piv_host + I.piv_host_off, I.piv_dev,
                                sizeof(int) * (size_t)I.ng, hipMemcpyDeviceToHost, s));
  }
  ISDF_HIP_TRY(hipEventRecord(pl.piv_ready, s));
  return 0;
}

int enqueue_back(Plan& pl, hipStream_t m, bool tune) {
  const int nb = pl.nb;
  hipEvent_t zfork = pl.ev[kEvZFork], zjoin = pl.ev[kEvZJoin];
  // The pivot rows of every instance, one launch.
  if (int rc = launch_gather_batched(pl.mode, pl.d_front, nb, pl.front_gather, m)) return rc;
  const int nvar = lp_variants(pl.mode);
  if (tune) {
    Timer tm;
    int best = 0;
    for (int v = 0; v < nvar; v++) {
      if (int rc = set_jobs(pl, v, -1)) return rc;
      if (int rc = tm.start(m)) return rc;
      if (int rc = launch_zct_lp_grouped(pl.mode, v, pl.d_zjobs, nb, pl.ztotal, pl.grid, m))
        return rc;
      if (int rc = tm.stop(m, &pl.zms[v])) return rc;
      if (pl.zms[v] < pl.zms[best]) best = v;
    }
    pl.zv = best;
  }
  if (tune)
    if (int rc = set_jobs(pl, pl.zv, 0)) return rc;
  // The pivot-row C C^T first, on the whole GPU (a few tens of microseconds); then
  // helper stream 0 runs the grouped Z C^T beside the rest of the Cholesky chain.
  if (int rc = launch_zct_lp_grouped(pl.mode, -1, pl.d_pjobs, nb, pl.ptotal, pl.grid, m))
    return rc;
  ISDF_HIP_TRY(hipEventRecord(zfork, m));
  if (!tune) {
    ISDF_HIP_TRY(hipStreamWaitEvent(pl.hs[0], zfork, 0));
    if (int rc = launch_zct_lp_grouped(pl.mode, pl.zv, pl.d_zjobs, nb, pl.ztotal, pl.grid,
                                       pl.hs[0]))
      return rc;
    ISDF_HIP_TRY(hipEventRecord(zjoin, pl.hs[0]));
  }
  // The Cholesky chain of every instance.
  if (int rc = launch_cct_lp_batched(pl.d_chol, nb, pl.max_rk, m)) return rc;
  if (int rc = launch_clear_info_batched(pl.d_chol, nb, m)) return rc;
  if (int rc = launch_potrf_batched(pl.d_chol, nb, pl.max_rk, m)) return rc;
  if (int rc = launch_make_fg_batched(pl.d_chol, nb, pl.max_rk, m)) return rc;
  if (int rc = launch_inverse_batched(pl.mode, pl.d_chol, nb, pl.max_rk, pl.max_ldz, m))
    return rc;
  if (!tune) ISDF_HIP_TRY(hipStreamWaitEvent(m, zjoin, 0));
  if (tune) {
    Timer tm;
    int best = 0;
    for (int v = 0; v < nvar; v++) {
      if (int rc = set_jobs(pl, pl.zv, v)) return rc;
      if (int rc = tm.start(m)) return rc;
      if (int rc = launch_theta_lp_grouped(pl.mode, v, pl.d_tjobs, nb, pl.ttotal, pl.grid, m))
        return rc;
      if (int rc = tm.stop(m, &pl.tms[v])) return rc;
      if (pl.tms[v] < pl.tms[best]) best = v;
    }
    pl.tv = best;
    if (int rc = set_jobs(pl, pl.zv, pl.tv)) return rc;
  } else {
    if (int rc = launch_theta_lp_grouped(pl.mode, pl.tv, pl.d_tjobs, nb, pl.ttotal, pl.grid, m))
      return rc;
  }
  hipLaunchKernelGGL(k_batch_status, dim3((nb + 63) / 64), dim3(64), 0, m, pl.d_src, nb,
                     pl.d_status);
  ISDF_HIP_LAUNCHED("k_batch_status");
  ISDF_HIP_TRY(hipMemcpyAsync(pl.status_host, pl.d_status, sizeof(InstStatus) * nb,
                              hipMemcpyDeviceToHost, m));
  return 0;
}

// piv of every instance from the pinned staging to the caller's buffers (an
// instance that falls back to the single solve overwrites its own afterwards).
void copy_piv_out(const Plan& pl, int* const* h_piv) {
  for (int i = 0; i < pl.nb; i++) {
    const Inst& I = pl.inst[i];
    std::memcpy(h_piv[i], pl.piv_host + I.piv_host_off, sizeof(int) * (size_t)I.ng);
  }
}

int solve_batched_locked(int nb, const isdf_problem* probs, const double* const* d_psi,
                         double* const* d_theta, int* const* h_piv, hipStream_t st,
                         isdf_report* reps, isdf_batch_report* brep) {
  if (nb <= 0 || !probs || !d_psi || !d_theta || !h_piv)
    return set_error(ISDF_EINVAL, "batched solve: empty batch or NULL argument");
  const int mode = probs[0].precision & ~ISDF_PREC_REFINE;
  if (mode != ISDF_PREC_FP32 && mode != ISDF_PREC_BF16X3)
    return set_error(ISDF_EINVAL,
                     "batched solve: the mixed forms (ISDF_PREC_FP32, ISDF_PREC_BF16X3) only");
  if (probs[0].precision & ISDF_PREC_REFINE)
    return set_error(ISDF_EINVAL, "batched solve: ISDF_PREC_REFINE is not supported");
  for (int i = 0; i < nb; i++) {
    long long ng = 0;
    if (int rc = validate_problem(&probs[i], &ng))
      return set_error(rc, "batch instance " + std::to_string(i) + ": " + isdf_hip_last_error());
    if (probs[i].select != ISDF_SELECT_KMEANS)
      return set_error(ISDF_EINVAL, "batched solve: K-means point selection only");
    if (probs[i].precision != probs[0].precision)
      return set_error(ISDF_EINVAL, "batched solve: every instance needs the same precision");
    if (probs[i].nv1 != probs[i].nv2 || probs[i].nc1 != probs[i].nc2)
      return set_error(ISDF_EINVAL, "batched solve: differing pair sets are not supported");
    if (!d_psi[i] || !d_theta[i] || !h_piv[i])
      return set_error(ISDF_EINVAL, "batch instance " + std::to_string(i) + ": NULL buffer");
  }
  int dev = 0;
  ISDF_HIP_TRY(hipGetDevice(&dev));
  Plan& pl = g_plan;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  ISDF_HIP_TRY(hipEventCreate(&t0));
  ISDF_HIP_TRY(hipEventCreate(&t1));
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } evg{t0, t1};
  int replayed = 0;
  if (!same_batch(pl, nb, probs, d_psi, d_theta, dev)) {
    if (int rc = build_plan(pl, nb, probs, d_psi, d_theta, dev)) {
      pl.release();
      pl.nb = 0;
      return rc;
    }
  }
  ISDF_HIP_TRY(hipEventRecord(t0, st));
  hipStream_t cps = pl.copy_stream;
  if (!pl.exec[0]) {
    // First call of this shape: eager run with tuning, then capture of both parts.
    if (int rc = enqueue_front(pl, st)) return rc;
    ISDF_HIP_TRY(hipEventRecord(pl.ev[kEvFork], st));
    if (int rc = issue_piv_copies(pl, cps, pl.ev[kEvFork])) return rc;
    if (int rc = enqueue_back(pl, st, true)) return rc;
    ISDF_HIP_TRY(hipEventRecord(t1, st));
    ISDF_HIP_TRY(hipStreamSynchronize(st));
    ISDF_HIP_TRY(hipStreamSynchronize(cps));
    for (int part = 0; part < 2; part++) {
      hipStream_t cs = nullptr;
      ISDF_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      ISDF_HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      int rc = part == 0 ? enqueue_front(pl, cs) : enqueue_back(pl, cs, false);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(cs, &g);
      (void)hipStreamDestroy(cs);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return set_error(ISDF_EHIP, std::string("graph capture failed: ") +
                                                           hipGetErrorString(e));
      pl.graph[part] = g;
      ISDF_HIP_TRY(hipGraphInstantiate(&pl.exec[part], g, nullptr, nullptr, 0));
    }
    copy_piv_out(pl, h_piv);
  } else {
    ISDF_HIP_TRY(hipGraphLaunch(pl.exec[0], st));
    ISDF_HIP_TRY(hipEventRecord(pl.ev[kEvFork], st));
    ISDF_HIP_TRY(hipGraphLaunch(pl.exec[1], st));
    ISDF_HIP_TRY(hipEventRecord(t1, st));
    // piv reaches pinned memory right after the K-means; it goes on to the caller's
    // buffers while the GPU runs the contractions and the Cholesky chains.
    if (int rc = issue_piv_copies(pl, cps, pl.ev[kEvFork])) return rc;
    ISDF_HIP_TRY(hipEventSynchronize(pl.piv_ready));
    copy_piv_out(pl, h_piv);
    ISDF_HIP_TRY(hipStreamSynchronize(st));
    replayed = 1;
  }
  float ms = 0.f;
  ISDF_HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
  int status = ISDF_OK;
  for (int i = 0; i < nb; i++) {
    const Inst& I = pl.inst[i];
    const InstStatus& s = pl.status_host[i];
    if (s.npt != I.ng) {
      // Not every grid point is a K-means point: the single solve handles it.
      if (int rc = solve_single_locked(&probs[i], d_psi[i], d_theta[i], h_piv[i], st,
                                       reps ? &reps[i] : nullptr, nullptr))
        return set_error(rc, "batch instance " + std::to_string(i) + ": " +
                                 isdf_hip_last_error());
      continue;
    }
    if (reps) {
      isdf_report& r = reps[i];
      std::memset(&r, 0, sizeof(r));
      r.rk = I.rk;
      r.rk_kmeans = s.rku;
      r.kmeans_iters = s.iters;
      r.npt = s.npt;
      r.potrf_info = s.info;
      r.eps = s.eps;
      r.ms_total = ms;
    }
    if (s.info != 0 && status == ISDF_OK)
      status = set_error(ISDF_ENOTSPD, "batch instance " + std::to_string(i) +
                                           ": Matrix is not positive definite.");
  }
  if (brep) {
    std::memset(brep, 0, sizeof(*brep));
    brep->ms_total = ms;
    brep->graph_replayed = replayed;
    brep->zct_tile = pl.zv;
    brep->theta_tile = pl.tv;
    for (int v = 0; v < 4; v++) {
      brep->zct_tile_ms[v] = pl.zms[v];
      brep->theta_tile_ms[v] = pl.tms[v];
    }
  }
  return status;
}

}  // namespace

void release_batch_plan() { g_plan.release(); g_plan.nb = 0; }

}  // namespace isdf_hip

extern "C" int isdf_hip_solve_batched(int nbatch, const isdf_problem* probs,
                                      const double* const* d_psi, double* const* d_theta,
                                      int* const* h_piv, void* stream, isdf_report* reps,
                                      isdf_batch_report* brep) {
  std::lock_guard<std::mutex> lk(isdf_hip::solve_mutex());
  isdf_hip::set_error(0, "");
  const int rc = isdf_hip::solve_batched_locked(nbatch, probs, d_psi, d_theta, h_piv,
                                                (hipStream_t)stream, reps, brep);
  if (rc == ISDF_EHIP) (void)hipDeviceSynchronize();  // drain what a failed call queued
  return rc;
}
