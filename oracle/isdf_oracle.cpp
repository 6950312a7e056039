// ISDF oracle: a single-rank CPU restatement of LibISDF's CPU ISDF
// ("scalapack" basis + "Kmeans" point selection).
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
// calls this file; it is the checker that tests/, the smoke entry and the CPU
// baseline leg of the benchmark compare the HIP path against.
//
// Every function cites the reference file:line it restates.  Floating-point
// operations follow the reference's order and rounding (the reference is built
// with g++ -O2 for x86-64 without FMA, so a*b+c is two roundings); this file is
// compiled with -ffp-contract=off for the same reason.  Where the reference
// calls MKL (dgemm, dpotrf, dtrsm) the restatement uses the netlib reference
// BLAS/LAPACK loop orders; MKL's own rounding is not reproducible, so those
// stages are pinned by tolerance and by mathematical invariants, not bitwise.
//
// OpenMP is used only where it cannot change a result: loops whose iterations
// write disjoint outputs with a fixed per-output operation order.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// rk rule, include/common/utility.hpp:18-23 (IRound) and src/common/isdf.cpp:37-41.
int orc_irank(int mu_points, int nv1, int nc1, int nv2, int nc2) {
  if (mu_points != 0) return mu_points;
  double a = std::sqrt((double)((nc1 + nv1) * (nc2 + nv2))) * 8.0;
  int b;
  if (a > 0) b = (a - (int)a < 0.5) ? (int)a : ((int)a + 1);
  else b = ((int)a - a < 0.5) ? (int)a : ((int)a - 1);
  return b;
}

// K-means weights, src/cpu/getpoints_Kmeans_MPI.cpp:33-45.
//   sumI(r) = sum_{i in [nocc-nv2, nocc+nc2)} psi_i(r)^2   (ascending i, from 0.0)
//   sumJ(r) = sum_{j in [nocc-nv1, nocc+nc1)} psi_j(r)^2
//   w(r)    = 0.0 + sumI(r)*sumJ(r)
void orc_weights(const double* psi, long ld, int ng, int nocc, int nv1, int nc1,
                 int nv2, int nc2, double* w) {
#pragma omp parallel for schedule(static)
  for (long kk = 0; kk < ng; kk++) {
    double sI = 0.0, sJ = 0.0;
    for (int i = nocc - nv2; i < nocc + nc2; i++) {
      double v = psi[(long)i * ld + kk];
      sI += v * v;
    }
    for (int j = nocc - nv1; j < nocc + nc1; j++) {
      double v = psi[(long)j * ld + kk];
      sJ += v * v;
    }
    double acc = 0.0;
    acc += sI * sJ;
    w[kk] = acc;
  }
}

// glibc rand() (TYPE_3 additive feedback generator, default seed 1) and
// libstdc++ std::random_shuffle, used only by the npt<rk branch of KMEAN
// (src/common/utility.cpp:667-681).  Restated from the published algorithm
// descriptions and pinned against the C library's rand() and libstdc++'s
// std::random_shuffle themselves (tests/test_oracle.py).  The one thing no
// file of the reference records is whether anything in its process draws
// from rand() before the shuffle; a fresh process (seed 1) is assumed.
struct GlibcRand {
  std::vector<int32_t> hist;
  explicit GlibcRand(uint32_t seed) {
    hist.resize(34);
    hist[0] = (int32_t)seed;
    for (int i = 1; i < 31; i++) {
      int64_t v = (16807LL * hist[i - 1]) % 2147483647LL;
      if (v < 0) v += 2147483647LL;
      hist[i] = (int32_t)v;
    }
    for (int i = 31; i < 34; i++) hist[i] = hist[i - 31];
    for (int i = 34; i < 344; i++)
      hist.push_back((int32_t)((uint32_t)hist[i - 31] + (uint32_t)hist[i - 3]));
  }
  int next() {
    size_t i = hist.size();
    int32_t v = (int32_t)((uint32_t)hist[i - 31] + (uint32_t)hist[i - 3]);
    hist.push_back(v);
    return (int)(((uint32_t)v) >> 1);
  }
};

// The first n values rand() returns after srand(seed) (glibc; seed 1 is the
// state of a process that never called srand).
void orc_glibc_rand(unsigned seed, int n, int* out) {
  GlibcRand g(seed);
  for (int i = 0; i < n; i++) out[i] = g.next();
}

// std::random_shuffle(first, first + len) of libstdc++ (bits/stl_algo.h): for
// i = 1 .. len-1, j = rand() % (i + 1), swap(first[i], first[j]) when i != j,
// with rand() in the state srand(seed) leaves.  utility.cpp:679 calls it on
// piv[npt..n) with the process's rand() state, seed 1 when nothing else drew.
void orc_random_shuffle(int* first, long len, unsigned seed) {
  GlibcRand rng(seed);
  for (long i = 1; i < len; i++) {
    long j = rng.next() % (i + 1);
    if (i != j) std::swap(first[i], first[j]);
  }
}

// Reference coordinate formula, src/common/utility.cpp:711-713 (points) and
// :758-760 (centroids).  Note x/dx0 need not equal the integer it came from;
// the restatement keeps every rounding of the original expression.
static inline void grid_coords(int tmp, const int nG[3], const double dx[3],
                               double* x, double* y, double* z) {
  double gx = (double)(tmp % nG[0]) * dx[0];
  double gy = ((double)(tmp % (nG[0] * nG[1])) - gx / dx[0]) / (double)nG[0] * dx[1];
  double gz = (((double)tmp - gx / dx[0]) - gy / dx[1] * (double)nG[0]) /
              (double)(nG[0] * nG[1]) * dx[2];
  *x = gx;
  *y = gy;
  *z = gz;
}

// pdist2 + findMin(Dim=1) fused: src/common/utility.cpp:598-615 and :535-596.
// D(i,c) = (d1*d1 + d2*d2) + d3*d3 ; label(i) = first c reaching the minimum.
// Four points go through one 4-wide vector at a time: every lane takes the
// same operations in the same order as the scalar loop (no FMA), so the labels
// are those of the scalar form (orc_set_scalar(1) selects it, for the tests).
typedef double v4d __attribute__((vector_size(32)));
typedef long long v4l __attribute__((vector_size(32)));
static int g_scalar = 0;

static inline int assign_one(double x, double y, double z, int rk, const double* cx,
                             const double* cy, const double* cz) {
  double best = 0.0;
  int bi = 0;
  for (int c = 0; c < rk; c++) {
    double d1 = x - cx[c];
    double d2 = y - cy[c];
    double d3 = z - cz[c];
    double d = d1 * d1 + d2 * d2 + d3 * d3;
    if (c == 0 || d < best) {
      best = d;
      bi = c;
    }
  }
  return bi;
}

static void assign_labels(int n, const double* X, const double* Y, const double* Z,
                          int rk, const double* cx, const double* cy, const double* cz,
                          int* label) {
  const int n4 = g_scalar ? 0 : n / 4;
#pragma omp parallel for schedule(static)
  for (int q = 0; q < n4; q++) {
    const int i = 4 * q;
    const v4d px = {X[i], X[i + 1], X[i + 2], X[i + 3]};
    const v4d py = {Y[i], Y[i + 1], Y[i + 2], Y[i + 3]};
    const v4d pz = {Z[i], Z[i + 1], Z[i + 2], Z[i + 3]};
    v4d d1 = px - cx[0], d2 = py - cy[0], d3 = pz - cz[0];
    v4d best = d1 * d1 + d2 * d2 + d3 * d3;
    v4l bi = {0, 0, 0, 0};
    for (int c = 1; c < rk; c++) {
      d1 = px - cx[c];
      d2 = py - cy[c];
      d3 = pz - cz[c];
      const v4d d = d1 * d1 + d2 * d2 + d3 * d3;
      const v4l lt = d < best;
      best = lt ? d : best;
      bi = lt ? (v4l){c, c, c, c} : bi;
    }
    for (int l = 0; l < 4; l++) label[i + l] = (int)bi[l];
  }
#pragma omp parallel for schedule(static)
  for (int i = 4 * n4; i < n; i++) label[i] = assign_one(X[i], Y[i], Z[i], rk, cx, cy, cz);
}

// findMin(Dim=0) for one centroid: the first point at the minimum distance.
// Lane l of the vector form scans the points i = l (mod 4); the lane minima
// are then combined by (distance, index), which is the first minimum overall.
static int nearest_point(int n, const double* X, const double* Y, const double* Z, double cx,
                         double cy, double cz) {
  double best = 0.0;
  int bi = 0;
  int i0 = 0;
  if (!g_scalar && n >= 4) {
    const int n4 = n / 4;
    v4d d1 = (v4d){X[0], X[1], X[2], X[3]} - cx;
    v4d d2 = (v4d){Y[0], Y[1], Y[2], Y[3]} - cy;
    v4d d3 = (v4d){Z[0], Z[1], Z[2], Z[3]} - cz;
    v4d vb = d1 * d1 + d2 * d2 + d3 * d3;
    v4l vi = {0, 1, 2, 3};
    for (int q = 1; q < n4; q++) {
      const int i = 4 * q;
      d1 = (v4d){X[i], X[i + 1], X[i + 2], X[i + 3]} - cx;
      d2 = (v4d){Y[i], Y[i + 1], Y[i + 2], Y[i + 3]} - cy;
      d3 = (v4d){Z[i], Z[i + 1], Z[i + 2], Z[i + 3]} - cz;
      const v4d d = d1 * d1 + d2 * d2 + d3 * d3;
      const v4l lt = d < vb;
      vb = lt ? d : vb;
      vi = lt ? (v4l){i, i + 1, i + 2, i + 3} : vi;
    }
    best = vb[0];
    bi = (int)vi[0];
    for (int l = 1; l < 4; l++)
      if (vb[l] < best || (vb[l] == best && vi[l] < bi)) {
        best = vb[l];
        bi = (int)vi[l];
      }
    i0 = 4 * n4;
  }
  for (int i = i0; i < n; i++) {
    double d1 = X[i] - cx;
    double d2 = Y[i] - cy;
    double d3 = Z[i] - cz;
    double d = d1 * d1 + d2 * d2 + d3 * d3;
    if (i == 0 || d < best) {
      best = d;
      bi = i;
    }
  }
  return bi;
}

void orc_set_scalar(int on) { g_scalar = on; }

// KMEAN, src/common/utility.cpp:631-988, at one MPI rank (the reductions,
// broadcasts and MINLOC all-reduce are identities there).
//
// In:  n = Ng, weight[n], rk, tolerances, iteration cap, domain length/grid.
// Out: piv[n] (must be zeroed by the caller, as src/common/isdf.cpp:43-44
//      does), *rk_out = number of distinct centers (the reference's local rk
//      after unique(), :970-972), *iters = iterations run, trace[s] = changed
//      label count of iteration s (the isdfOFS trace, :920), labels[n] = final
//      labels (may be NULL), cent[3*rk] = final centroids x|y|z, blocks of rk
//      with the first rk' (after compaction) entries written (may be NULL),
//      *npt_out = points above the density threshold.
// Returns 0 on the normal path, 1 on the npt<rk branch (:667-681).
int orc_kmeans(int n, const double* weight, int rk, double kmeans_tol, int max_iter,
               double df_tol, const double length[3], const int ngrid[3], int* piv,
               int* rk_out, int* iters, int* trace, int* labels, double* cent,
               int* npt_out) {
  const int rk0 = rk;
  int npt;
  std::vector<int> index(n, 0);  // value-initialised; entries past npt stay 0
  double maxW = 0.0;
  if (df_tol > 1e-16) {
    for (int i = 0; i < n; i++)
      if (weight[i] > maxW) maxW = weight[i];  // findMax, numvec_impl.hpp:259-266
    npt = 0;
    for (int i = 0; i < n; i++) {
      if (weight[i] > df_tol * maxW) {
        index[npt] = i;
        npt++;
      }
    }
  } else {
    npt = n;
    for (int i = 0; i < n; i++) index[i] = i;
  }
  if (npt_out) *npt_out = npt;

  if (npt < rk) {  // :667-681
    int k0 = 0, k1 = 0;
    for (int i = 0; i < npt; i++) {
      if (i == index[k0]) {
        piv[k0] = i;
        k0 = std::min(k0 + 1, rk - 1);
      } else {
        piv[npt + k1] = i;
        k1++;
      }
    }
    orc_random_shuffle(piv + npt, (long)n - npt, 1u);
    if (rk_out) *rk_out = rk;
    if (iters) *iters = 0;
    return 1;
  }

  // One rank: nptLocal = n, indexSta = 0 (:682-690).  Reading index[i] for
  // i >= npt reads the zero-initialised tail of the vector (grid point 0).
  double dx[3];
  int nG[3];
  for (int d = 0; d < 3; d++) {
    nG[d] = ngrid[d];
    dx[d] = length[d] / nG[d];
  }
  std::vector<int> indexLocal(n);
  std::vector<double> X(n), Y(n), Z(n), WL(n);
  for (int i = 0; i < n; i++) {
    int tmp = index[i];
    indexLocal[i] = tmp;
    WL[i] = weight[tmp];
    grid_coords(tmp, nG, dx, &X[i], &Y[i], &Z[i]);
  }

  // Initial centroids: piv is zeroed, so piv[0]==piv[1] and Cinit = the first
  // rk entries of index (:725-753).
  std::vector<double> cx(rk), cy(rk), cz(rk);
  for (int i = 0; i < rk; i++) grid_coords(index[i], nG, dx, &cx[i], &cy[i], &cz[i]);

  std::vector<int> label(n), last(n);
  assign_labels(n, X.data(), Y.data(), Z.data(), rk, cx.data(), cy.data(), cz.data(),
                label.data());

  int s = 0;
  int flag = n;
  double maxF = kmeans_tol * n;
  std::vector<double> cnt, sx, sy, sz;
  while (flag > maxF && s < max_iter) {  // :813
    cnt.assign(rk, 0.0);
    sx.assign(rk, 0.0);
    sy.assign(rk, 0.0);
    sz.assign(rk, 0.0);
    for (int i = 0; i < n; i++) {  // :816-822
      int l = label[i];
      cnt[l] += WL[i];
      sx[l] += WL[i] * X[i];
      sy[l] += WL[i] * Y[i];
      sz[l] += WL[i] * Z[i];
    }
    // :831-854: divide, dropping empty clusters by swap-with-last.
    for (int i = 0; i < rk; i++) {
      if (cnt[i] != 0.0) {
        sx[i] = sx[i] / cnt[i];
        sy[i] = sy[i] / cnt[i];
        sz[i] = sz[i] / cnt[i];
      } else {
        rk--;
        sx[i] = sx[rk];
        sy[i] = sy[rk];
        sz[i] = sz[rk];
        cnt[i] = cnt[rk];
        i--;
      }
    }
    cx.assign(sx.begin(), sx.begin() + rk);
    cy.assign(sy.begin(), sy.begin() + rk);
    cz.assign(sz.begin(), sz.begin() + rk);
    last.swap(label);
    assign_labels(n, X.data(), Y.data(), Z.data(), rk, cx.data(), cy.data(), cz.data(),
                  label.data());
    flag = 0;
    for (int i = 0; i < n; i++)
      if (last[i] != label[i]) flag++;
    if (trace) trace[s] = flag;
    s++;
  }
  if (iters) *iters = s;

  // findMin(Dim=0): per centroid the first point at minimum distance (:926-945).
  std::vector<int> imin(rk);
#pragma omp parallel for schedule(dynamic, 8)
  for (int c = 0; c < rk; c++)
    imin[c] = indexLocal[nearest_point(n, X.data(), Y.data(), Z.data(), cx[c], cy[c], cz[c])];
  // unique (:617-629): sort + dedupe.
  std::sort(imin.begin(), imin.end());
  imin.erase(std::unique(imin.begin(), imin.end()), imin.end());
  int rku = (int)imin.size();
  // piv fill (:973-983).
  int k0 = 0, k1 = 0;
  for (int i = 0; i < n; i++) {
    if (i == imin[k0]) {
      piv[k0] = i;
      k0 = std::min(k0 + 1, rku - 1);
    } else {
      piv[rku + k1] = i;
      k1++;
    }
  }
  if (rk_out) *rk_out = rku;
  if (labels) std::memcpy(labels, label.data(), sizeof(int) * n);
  if (cent) {  // stride: the caller's rk; entries past the compacted count stay as given
    for (int c = 0; c < rk; c++) {
      cent[c] = cx[c];
      cent[rk0 + c] = cy[c];
      cent[2 * rk0 + c] = cz[c];
    }
  }
  return 0;
}

// Psi(mu) gather, src/cpu/getbasis_scalapack.cpp:249-269:
//   psimu(k, mu) = psi(piv[mu], k), k < nstate, mu < rk   (psimu ld = nstate)
void orc_gather(const double* psi, long ld, int nstate, const int* piv, int rk,
                double* psimu) {
  for (int mu = 0; mu < rk; mu++)
    for (int k = 0; k < nstate; k++) psimu[(long)mu * nstate + k] = psi[(long)k * ld + piv[mu]];
}

// Z C^T, src/cpu/getbasis_scalapack.cpp:297-369.
//   P1 = psi[:, s1:s1+K1] * psimu[s1:s1+K1, :]   (my_pdgemm N,N at 1-based column nocc-nv1+1)
//   equal pair sets:  out = P1 .* P1            (:360-367)
//   otherwise:        out = P1 .* P2, P2 from S2 (:319-349)
// The dgemm restatement sums k ascending from 0.0 with separate roundings.
// Rows [r0, r1) only, so tests can sample rows of a large problem.
void orc_zct_rows(const double* psi, long ld, int nstate, const double* psimu, int rk,
                  int nocc, int nv1, int nc1, int nv2, int nc2, long r0, long r1,
                  double* out, long ldo) {
  int s1 = nocc - nv1, K1 = nv1 + nc1;
  int s2 = nocc - nv2, K2 = nv2 + nc2;
  bool same = !(nv1 != nv2 || nc1 != nc2);
#pragma omp parallel for schedule(static)
  for (long r = r0; r < r1; r++) {
    std::vector<double> a1(K1), a2(K2);
    for (int k = 0; k < K1; k++) a1[k] = psi[(long)(s1 + k) * ld + r];
    for (int k = 0; k < K2; k++) a2[k] = psi[(long)(s2 + k) * ld + r];
    for (int mu = 0; mu < rk; mu++) {
      const double* b = psimu + (long)mu * nstate;
      double p1 = 0.0;
      for (int k = 0; k < K1; k++) p1 += a1[k] * b[s1 + k];
      double v;
      if (same) {
        v = p1 * p1;
      } else {
        double p2 = 0.0;
        for (int k = 0; k < K2; k++) p2 += a2[k] * b[s2 + k];
        v = p1 * p2;
      }
      out[(long)mu * ldo + (r - r0)] = v;
    }
  }
}

// The same Z C^T values for an arbitrary list of rows (sampling at full size).
void orc_zct_pick(const double* psi, long ld, int nstate, const double* psimu, int rk, int nocc,
                  int nv1, int nc1, int nv2, int nc2, const int* rows, long nrows, double* out,
                  long ldo) {
  int s1 = nocc - nv1, K1 = nv1 + nc1;
  int s2 = nocc - nv2, K2 = nv2 + nc2;
  bool same = !(nv1 != nv2 || nc1 != nc2);
#pragma omp parallel for schedule(static)
  for (long q = 0; q < nrows; q++) {
    long r = rows[q];
    for (int mu = 0; mu < rk; mu++) {
      const double* b = psimu + (long)mu * nstate;
      double p1 = 0.0;
      for (int k = 0; k < K1; k++) p1 += psi[(long)(s1 + k) * ld + r] * b[s1 + k];
      double v;
      if (same) {
        v = p1 * p1;
      } else {
        double p2 = 0.0;
        for (int k = 0; k < K2; k++) p2 += psi[(long)(s2 + k) * ld + r] * b[s2 + k];
        v = p1 * p2;
      }
      out[(long)mu * ldo + q] = v;
    }
  }
}

// C C^T extraction and regularisation, src/cpu/getbasis_scalapack.cpp:390-428:
//   P(mu,nu) = Xi(piv[mu], nu); trace = sum_nu P(nu,nu) ascending;
//   eps = 1e-10 * trace / rk; P(nu,nu) += eps.  Returns eps.
double orc_cct(const double* xi, long ldx, const int* piv, int rk, double* P) {
  for (int nu = 0; nu < rk; nu++)
    for (int mu = 0; mu < rk; mu++) P[(long)nu * rk + mu] = xi[(long)nu * ldx + piv[mu]];
  double trace = 0.0;
  for (int nu = 0; nu < rk; nu++) trace += P[(long)nu * rk + nu];
  double eps = 1e-10 * trace / (double)rk;
  for (int nu = 0; nu < rk; nu++) P[(long)nu * rk + nu] += eps;
  return eps;
}

// Lower Cholesky, the netlib dpotf2 loop order (the reference calls pdpotrf
// 'L', src/cpu/getbasis_scalapack.cpp:455; lapack::Potrf, src/cpu/lapack.cpp:52-68).
// Returns LAPACK's info: 0, or j+1 for the first non-positive pivot.  The strict
// upper triangle is left untouched.
int orc_potrf(double* A, int n) {
  for (int j = 0; j < n; j++) {
    double ajj = A[(long)j * n + j];
    for (int k = 0; k < j; k++) ajj -= A[(long)k * n + j] * A[(long)k * n + j];
    if (!(ajj > 0.0)) {
      A[(long)j * n + j] = ajj;
      return j + 1;
    }
    ajj = std::sqrt(ajj);
    A[(long)j * n + j] = ajj;
    double r = 1.0 / ajj;
#pragma omp parallel for schedule(static)
    for (int i = j + 1; i < n; i++) {
      double v = A[(long)j * n + i];
      for (int k = 0; k < j; k++) v -= A[(long)k * n + i] * A[(long)k * n + j];
      A[(long)j * n + i] = v * r;
    }
  }
  return 0;
}

// potrs as the reference does it: two right-side triangular solves in place,
// src/cpu/getbasis_scalapack.cpp:469-478 (pdtrsm R,L,T,N then R,L,N,N).
// Netlib dtrsm loop orders (reciprocal of the diagonal, then AXPY updates).
// B is rows x n, column-major with leading dimension ldb.
void orc_potrs_rows(const double* L, int n, double* B, long ldb, long rows) {
  std::vector<double> rdiag(n);
  for (int k = 0; k < n; k++) rdiag[k] = 1.0 / L[(long)k * n + k];
#pragma omp parallel for schedule(static)
  for (long i = 0; i < rows; i++) {
    // B := B * inv(L^T): for k ascending, B(:,k) *= 1/L(k,k); B(:,j) -= L(j,k) B(:,k), j > k.
    for (int k = 0; k < n; k++) {
      double bk = rdiag[k] * B[(long)k * ldb + i];
      B[(long)k * ldb + i] = bk;
      for (int j = k + 1; j < n; j++) {
        double t = L[(long)k * n + j];
        if (t != 0.0) B[(long)j * ldb + i] -= t * bk;
      }
    }
    // B := B * inv(L): for k descending, B(:,k) -= sum_{j>k} L(j,k) B(:,j); B(:,k) *= 1/L(k,k).
    for (int k = n - 1; k >= 0; k--) {
      double v = B[(long)k * ldb + i];
      for (int j = k + 1; j < n; j++) {
        double t = L[(long)k * n + j];
        if (t != 0.0) v -= t * B[(long)j * ldb + i];
      }
      B[(long)k * ldb + i] = rdiag[k] * v;
    }
  }
}

// Fit error of src/common/checkcode.cpp:62-166:
//   Z(r, (k-s1)*K2 + (j-s2)) = psi_k(r) psi_j(r)
//   C(mu, same column)      = psi_k(r_piv[mu]) psi_j(r_piv[mu])
//   err = sqrt(sum (Z - Theta C)^2) / sqrt(sum Z^2)
// Streams over pair columns so that Z is never materialised (Ng*K1*K2 doubles).
// Returns err; *zt_norm receives sqrt(sum Z^2).
double orc_fit_error(const double* psi, long ld, int ng, int nstate, int nocc, int nv1,
                     int nc1, int nv2, int nc2, const double* theta, long ldt, int rk,
                     const int* piv, double* z_norm) {
  int s1 = nocc - nv1, K1 = nv1 + nc1;
  int s2 = nocc - nv2, K2 = nv2 + nc2;
  long npair = (long)K1 * K2;
  double err = 0.0, nrm = 0.0;
  const int blk = 32;
  std::vector<double> cmat((size_t)rk * blk);
  for (long p0 = 0; p0 < npair; p0 += blk) {
    int nb = (int)std::min<long>(blk, npair - p0);
    for (int b = 0; b < nb; b++) {
      long p = p0 + b;
      int k = (int)(p / K2), j = (int)(p % K2);
      for (int mu = 0; mu < rk; mu++)
        cmat[(size_t)b * rk + mu] =
            psi[(long)(s1 + k) * ld + piv[mu]] * psi[(long)(s2 + j) * ld + piv[mu]];
    }
    double e = 0.0, z = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : e, z)
    for (long r = 0; r < ng; r++) {
      double acc[32];
      for (int b = 0; b < nb; b++) acc[b] = 0.0;
      for (int mu = 0; mu < rk; mu++) {
        double t = theta[(long)mu * ldt + r];
        for (int b = 0; b < nb; b++) acc[b] += t * cmat[(size_t)b * rk + mu];
      }
      for (int b = 0; b < nb; b++) {
        long p = p0 + b;
        int k = (int)(p / K2), j = (int)(p % K2);
        double zz = psi[(long)(s1 + k) * ld + r] * psi[(long)(s2 + j) * ld + r];
        double d = zz - acc[b];
        e += d * d;
        z += zz * zz;
      }
    }
    err += e;
    nrm += z;
  }
  if (z_norm) *z_norm = std::sqrt(nrm);
  return std::sqrt(err) / std::sqrt(nrm);
}

// Whole CPU ISDF at one rank, src/common/isdf.cpp:8-127 with s="scalapack",
// s1="Kmeans".  psi is Ng x nstate column-major (ld >= Ng); theta receives
// Ng x rk (ld Ng).  piv[Ng] is zeroed here as the reference does.
// Returns 0, or 2 when the Cholesky fails ("Matrix is not positive definite.").
int orc_isdf(const double* psi, long ld, const double length[3], const int ngrid[3],
             int nocc, int nstate, int nv1, int nc1, int nv2, int nc2, int mu_points,
             double* theta, int* piv, int max_iter, int* rk_kmeans, int* iters,
             int* trace) {
  long ng = (long)ngrid[0] * ngrid[1] * ngrid[2];
  int rk = orc_irank(mu_points, nv1, nc1, nv2, nc2);
  std::memset(piv, 0, sizeof(int) * ng);
  std::vector<double> w(ng);
  orc_weights(psi, ld, (int)ng, nocc, nv1, nc1, nv2, nc2, w.data());
  int npt = 0;
  orc_kmeans((int)ng, w.data(), rk, 1e-5, max_iter, 1e-8, length, ngrid, piv, rk_kmeans,
             iters, trace, nullptr, nullptr, &npt);
  std::vector<double> psimu((size_t)nstate * rk);
  orc_gather(psi, ld, nstate, piv, rk, psimu.data());
  orc_zct_rows(psi, ld, nstate, psimu.data(), rk, nocc, nv1, nc1, nv2, nc2, 0, ng, theta, ng);
  std::vector<double> P((size_t)rk * rk);
  orc_cct(theta, ng, piv, rk, P.data());
  int info = orc_potrf(P.data(), rk);
  if (info != 0) return 2;
  orc_potrs_rows(P.data(), rk, theta, ng, ng);
  return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
