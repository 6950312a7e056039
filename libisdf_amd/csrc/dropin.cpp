// C++ drop-in isdf::isdffunc::ISDF over the C ABI (include/isdf_hip.h).
//
// The device-pointer overload mirrors the GPU driver of LibISDF,
// src/common/isdf.cpp:132-209: rank rule, piv zeroing, the solve, the isdfOFS
// log lines, and the check=true validator (src/common/checkcode.cpp, reached
// from isdf.cpp:193-206), which runs on the device (isdf_hip_check_device).  Unlike the reference there is no row->column
// relayout: at one rank that relayout is the identity, so Theta is solved
// directly in d_thetaCol.  The host-pointer overload mirrors the CPU driver,
// src/common/isdf.cpp:8-127, with its s / s1 dispatch, on the same GPU solve.

#define ISDF_DECLARE_BOTH_INTERFACES 1

#include <hip/hip_runtime.h>

#include <cmath>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/isdf.hpp"
#include "../../include/isdf_hip.h"

namespace isdf {

std::ofstream isdfOFS;

void ErrorHandling(const char* msg) {
  isdfOFS << std::endl << "ERROR!" << std::endl << msg << std::endl << std::endl;
  throw std::runtime_error(msg);
}

namespace isdffunc {

namespace {

isdf_problem make_problem(const Domain& domain, int nocc, int nstate, int nv1, int nc1, int nv2,
                          int nc2, int mu_points, int KmeansMaxIter_ISDF) {
#ifdef ISDF_WITH_MPI
  int mpisize = 1;
  MPI_Comm_size(domain.comm, &mpisize);
  if (mpisize != 1) ErrorHandling("libisdf_hip runs one rank; the communicator has more.");
#endif
  isdf_problem p{};
  for (int d = 0; d < 3; d++) {
    p.ngrid[d] = domain.numGrid[d];
    p.length[d] = domain.length[d];
  }
  p.nocc = nocc;
  p.nstate = nstate;
  p.nv1 = nv1;
  p.nc1 = nc1;
  p.nv2 = nv2;
  p.nc2 = nc2;
  p.mu_points = mu_points;
  p.kmeans_max_iter = KmeansMaxIter_ISDF;
  p.precision = ISDF_PREC_FP64;
  p.ld_psi = 0;
  return p;
}

// The K-means trace as utility.cpp:920-923 prints it, then the stage times.
void log_solve(const isdf_report& rep, const std::vector<int>& trace) {
  for (int s = 0; s < rep.kmeans_iters; s++) isdfOFS << trace[(size_t)s] << " ";
  isdfOFS << std::endl << "Converged in " << rep.kmeans_iters << " iterations." << std::endl;
  isdfOFS << "Time for pivQR_ with Kmeans      = " << (rep.ms_weights + rep.ms_kmeans + rep.ms_select) * 1e-3
          << " [s]" << std::endl;
  isdfOFS << "Time for Z*C^T                   = " << rep.ms_zct * 1e-3 << " [s]" << std::endl;
  isdfOFS << "Time for C*C^T                   = " << rep.ms_cct * 1e-3 << " [s]" << std::endl;
  isdfOFS << "Time for Potrf                   = " << rep.ms_potrf * 1e-3 << " [s]" << std::endl;
  isdfOFS << "Time for Trsm                    = " << rep.ms_potrs * 1e-3 << " [s]" << std::endl;
  isdfOFS << "Time for ISDF time =" << rep.ms_total * 1e-3 << " [s]" << std::endl;
}

// The error lines of checkcode.cpp:164-166 and :310-312.
void log_check(const isdf_check& chk) {
  isdfOFS << "sqrt((rhoij-rho^hat)^2) =" << chk.fit_error * chk.z_norm << std::endl;
  isdfOFS << "sqrt(rhoij^2)= " << chk.z_norm << std::endl;
  isdfOFS << "sqrt((rhoij-rho^hat)^2)/sqrt(rhoij^2)= " << chk.fit_error << std::endl;
  isdfOFS << "sqrt((ijkl-ijkl^hat)^2)) =" << chk.coulomb_error * chk.coulomb_norm << std::endl;
  isdfOFS << "sqrt(ijkl^2) =" << chk.coulomb_norm << std::endl;
  isdfOFS << "sqrt((ijkl-ijkl^hat)^2)/sqrt(ijkl^2) =" << chk.coulomb_error << std::endl;
}

}  // namespace

void ISDF(double* d_psirow, Domain domain, int nocc, int nstate, int nv1, int nc1, int nv2,
          int nc2, int mu_points, double* d_thetaCol, int* piv, int KmeansMaxIter_ISDF,
          bool check) {
  const isdf_problem p =
      make_problem(domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, KmeansMaxIter_ISDF);
  const int rk = isdf_hip_rank(&p);
  isdfOFS << "rk=" << rk << std::endl;

  isdf_report rep{};
  std::vector<int> trace((size_t)(KmeansMaxIter_ISDF > 0 ? KmeansMaxIter_ISDF : 1));
  isdf_debug dbg{};
  dbg.h_trace = trace.data();
  int rc = isdf_hip_solve(&p, d_psirow, d_thetaCol, piv, nullptr, &rep, &dbg);
  if (rc != ISDF_OK) ErrorHandling(isdf_hip_last_error());
  log_solve(rep, trace);

  if (check) {
    // Both halves of checkcode on the device, on the caller's buffers.
    isdf_check chk{};
    rc = isdf_hip_check_device(&p, d_psirow, d_thetaCol, piv, 1, nullptr, &chk);
    if (rc != ISDF_OK) ErrorHandling(isdf_hip_last_error());
    log_check(chk);
  }
}

void ISDF(double* psirow, Domain domain, int nocc, int nstate, int nv1, int nc1, int nv2,
          int nc2, int mu_points, double* thetaCol, int* piv, int KmeansMaxIter_ISDF,
          int scalblocksize, std::string s, std::string s1, bool check) {
  const isdf_problem p =
      make_problem(domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, KmeansMaxIter_ISDF);
  const int rk = isdf_hip_rank(&p);
  isdfOFS << "rk=" << rk << std::endl;
  // isdf.cpp:43-44 zeroes piv before the backend dispatch.
  const long long ng = (long long)p.ngrid[0] * p.ngrid[1] * p.ngrid[2];
  if (piv)
    for (long long g = 0; g < ng; g++) piv[g] = 0;
  // isdf.cpp:60 and :98: "MPI" and "scalapack" differ only in how the basis
  // is distributed over ranks; any other s computes nothing.
  if (s != "MPI" && s != "scalapack") return;
  if (s1 != "Kmeans" && s1 != "QRCP") ErrorHandling("Wrong select points method");
  if (scalblocksize <= 0) ErrorHandling("scalblocksize must be positive.");
  isdf_problem q = p;
  q.select = s1 == "QRCP" ? ISDF_SELECT_QRCP : ISDF_SELECT_KMEANS;

  isdf_report rep{};
  std::vector<int> trace((size_t)(KmeansMaxIter_ISDF > 0 ? KmeansMaxIter_ISDF : 1));
  isdf_debug dbg{};
  dbg.h_trace = trace.data();
  int rc = isdf_hip_solve_host(&q, psirow, thetaCol, piv, nullptr, &rep, &dbg);
  if (rc != ISDF_OK) ErrorHandling(isdf_hip_last_error());
  log_solve(rep, trace);
  if (check) {
    isdf_check chk{};
    rc = isdf_hip_check(&q, psirow, thetaCol, piv, 1, &chk);
    if (rc != ISDF_OK) ErrorHandling(isdf_hip_last_error());
    log_check(chk);
  }
}

}  // namespace isdffunc
}  // namespace isdf
