// isdf.hpp — source-compatible drop-in for LibISDF's ISDF entry points.
//
// Replaces include/isdf.hpp of LibISDF.  The GPU build (-DGPU) sees
//   include/isdf.hpp:37  void isdf::isdffunc::ISDF(double *d_psirow, Domain domain,
//                          int nocc, int nstate, int nv1, int nc1, int nv2, int nc2,
//                          int mu_points, double *d_thetaCol, int *piv,
//                          int KmeansMaxIter_ISDF, bool check=false);
// and any other build the host-buffer interface of include/isdf.hpp:39, both
// computed on the GPU by this library.  The header also provides the Domain
// type (include/common/domain.hpp:8-36), the isdfOFS log stream and
// ErrorHandling (src/common/global.cpp:5-11).
//
// Behaviour follows the reference at one rank: d_psirow is the caller's device
// Ng x nstate column-major matrix, d_thetaCol the caller's device Ng x rk
// matrix (column mu belongs to grid point piv[mu]), piv the caller's host
// int[Ng]; the host interface takes the same matrices in host memory.  The
// call blocks until the result is in place.  Errors are logged to
// isdfOFS and thrown as std::runtime_error, as ErrorHandling does.
//
// Communicators: build with ISDF_WITH_MPI (and mpi.h on the include path) to
// keep MPI_Comm members; the library runs one rank and throws if the
// communicator holds more than one process.  Without it the members are
// opaque integers.
#pragma once

#include <fstream>
#include <stdexcept>
#include <string>

#ifdef ISDF_WITH_MPI
#include <mpi.h>
#endif

namespace isdf {

typedef int Int;
typedef double Real;

#ifdef ISDF_WITH_MPI
typedef MPI_Comm Comm;
#else
typedef int Comm;  // opaque; 0 stands for the single-process world
#endif

// Minimal Point3 / Index3 with the accessors Domain users need
// (LibISDF's tinyvec: operator[] / operator()).
template <class F>
struct Vec3 {
  F v[3];
  Vec3() : v{F(0), F(0), F(0)} {}
  Vec3(F a, F b, F c) : v{a, b, c} {}
  F& operator[](int i) { return v[i]; }
  const F& operator[](int i) const { return v[i]; }
  F& operator()(int i) { return v[i]; }
  const F& operator()(int i) const { return v[i]; }
  F* Data() { return v; }
  const F* Data() const { return v; }
};
typedef Vec3<Real> Point3;
typedef Vec3<Int> Index3;

struct Domain {
  Point3 length;
  Point3 posStart;
  Index3 numGrid;
  Index3 numGridFine;
  Comm comm;
  Comm rowComm;
  Comm colComm;

  Domain()
      : length(0.0, 0.0, 0.0),
        posStart(0.0, 0.0, 0.0),
        numGrid(0, 0, 0),
        numGridFine(0, 0, 0),
#ifdef ISDF_WITH_MPI
        comm(MPI_COMM_WORLD),
        rowComm(MPI_COMM_WORLD),
        colComm(MPI_COMM_WORLD)
#else
        comm(0),
        rowComm(0),
        colComm(0)
#endif
  {
  }

  Real Volume() const { return length[0] * length[1] * length[2]; }
  Int NumGridTotal() const { return numGrid[0] * numGrid[1] * numGrid[2]; }
  Int NumGridTotalFine() const { return numGridFine[0] * numGridFine[1] * numGridFine[2]; }
};

// Log stream opened by the caller (the reference opens "isdf.<rank>").
extern std::ofstream isdfOFS;

// Logs msg to isdfOFS and throws std::runtime_error(msg).
void ErrorHandling(const char* msg);

namespace isdffunc {

// As in the reference header, a -DGPU build sees the device-pointer ISDF and
// any other build the host-pointer one; the library defines both
// (ISDF_DECLARE_BOTH_INTERFACES).
#if defined(GPU) || defined(ISDF_DECLARE_BOTH_INTERFACES)
void ISDF(double* d_psirow, Domain domain, int nocc, int nstate, int nv1, int nc1, int nv2,
          int nc2, int mu_points, double* d_thetaCol, int* piv, int KmeansMaxIter_ISDF,
          bool check = false);
#endif

#if !defined(GPU) || defined(ISDF_DECLARE_BOTH_INTERFACES)
// include/isdf.hpp:39 of the reference: host buffers (psirow Ng x nstate,
// thetaCol Ng x rk, piv Ng; column-major, one rank), computed on the GPU.
// s = "scalapack" or "MPI" (the same Theta at one rank); any other s only
// zeroes piv, as the reference does.  s1 = "Kmeans" or "QRCP" (pivoted QR of
// the Gaussian-sketched pair products, getpoints_QRCP.cpp); anything else
// throws "Wrong select points method".  scalblocksize
// (the ScaLAPACK block size) has no role at one rank and must be positive.
void ISDF(double* psirow, Domain domain, int nocc, int nstate, int nv1, int nc1, int nv2,
          int nc2, int mu_points, double* thetaCol, int* piv, int KmeansMaxIter_ISDF,
          int scalblocksize = 64, std::string s = "scalapack", std::string s1 = "Kmeans",
          bool check = false);
#endif

}  // namespace isdffunc
}  // namespace isdf
