// A LibISDF caller written against include/isdf.hpp, in the pattern of the
// reference's examples/si8/readpsi.cpp:170-240: open isdfOFS, allocate the
// buffers, call isdffunc::ISDF, read Theta back.  Built twice: with -DGPU it
// calls the device-pointer ISDF (include/isdf.hpp:37), without it the
// host-pointer one (include/isdf.hpp:39).  Test infrastructure for the
// drop-in tests; not part of the library.
//
// usage: caller psi.bin nx ny nz lx ly lz nocc nstate nv1 nc1 nv2 nc2 mu maxiter check
//               theta.bin piv.bin log [s s1 blocksize]
// psi.bin: Ng x nstate doubles, column-major.  Writes Ng x rk doubles to
// theta.bin and Ng ints to piv.bin.  Exit status 0 on success, 2 when ISDF
// threw (its message is in the log, as ErrorHandling writes it).

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#if defined(GPU)
#include <hip/hip_runtime.h>
#endif

#include "isdf.hpp"

using namespace isdf;

static std::vector<char> read_file(const char* path) {
  std::vector<char> b;
  FILE* f = std::fopen(path, "rb");
  if (!f) return b;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) b.insert(b.end(), buf, buf + n);
  std::fclose(f);
  return b;
}

static bool write_file(const char* path, const void* p, size_t bytes) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  const bool ok = std::fwrite(p, 1, bytes, f) == bytes;
  std::fclose(f);
  return ok;
}

int main(int argc, char** argv) {
  if (argc < 20) {
    std::fprintf(stderr, "usage: %s psi.bin nx ny nz lx ly lz nocc nstate nv1 nc1 nv2 nc2 mu "
                         "maxiter check theta.bin piv.bin log [s s1 blocksize]\n", argv[0]);
    return 1;
  }
  Domain domain_;
  for (int d = 0; d < 3; d++) {
    domain_.numGrid[d] = std::atoi(argv[2 + d]);
    domain_.length[d] = std::atof(argv[5 + d]);
  }
  const int nocc = std::atoi(argv[8]), nstate = std::atoi(argv[9]);
  const int Nv1 = std::atoi(argv[10]), Nc1 = std::atoi(argv[11]);
  const int Nv2 = std::atoi(argv[12]), Nc2 = std::atoi(argv[13]);
  const int Nu = std::atoi(argv[14]), KmeansMaxIter_ISDF = std::atoi(argv[15]);
  const bool check = std::atoi(argv[16]) != 0;
  const long long Nr = domain_.NumGridTotal();
  const int rk = Nu;

  std::vector<char> raw = read_file(argv[1]);
  if (raw.size() != sizeof(double) * (size_t)(Nr * nstate)) {
    std::fprintf(stderr, "psi file holds %zu bytes, expected %zu\n", raw.size(),
                 sizeof(double) * (size_t)(Nr * nstate));
    return 1;
  }
  const double* psirow = (const double*)raw.data();
  std::vector<double> theta((size_t)(Nr * rk), 0.0);
  std::vector<int> piv((size_t)Nr, -1);
  isdfOFS.open(argv[19]);

  int status = 0;
  try {
#if defined(GPU)
    double *d_psirow = nullptr, *d_theta = nullptr;
    if (hipMalloc((void**)&d_psirow, raw.size()) != hipSuccess ||
        hipMalloc((void**)&d_theta, sizeof(double) * theta.size()) != hipSuccess ||
        hipMemcpy(d_psirow, psirow, raw.size(), hipMemcpyHostToDevice) != hipSuccess) {
      std::fprintf(stderr, "device allocation or copy failed\n");
      return 1;
    }
    isdffunc::ISDF(d_psirow, domain_, nocc, nstate, Nv1, Nc1, Nv2, Nc2, Nu, d_theta, piv.data(),
                   KmeansMaxIter_ISDF, check);
    if (hipMemcpy(theta.data(), d_theta, sizeof(double) * theta.size(), hipMemcpyDeviceToHost) !=
        hipSuccess) {
      std::fprintf(stderr, "device copy failed\n");
      return 1;
    }
    (void)hipFree(d_psirow);
    (void)hipFree(d_theta);
#else
    const std::string s = argc > 20 ? argv[20] : "scalapack";
    const std::string s1 = argc > 21 ? argv[21] : "Kmeans";
    const int blocksize = argc > 22 ? std::atoi(argv[22]) : 32;
    isdffunc::ISDF((double*)psirow, domain_, nocc, nstate, Nv1, Nc1, Nv2, Nc2, Nu, theta.data(),
                   piv.data(), KmeansMaxIter_ISDF, blocksize, s, s1, check);
#endif
  } catch (const std::runtime_error& e) {
    std::fprintf(stderr, "ISDF threw: %s\n", e.what());
    status = 2;
  }
  isdfOFS.close();
  if (!write_file(argv[17], theta.data(), sizeof(double) * theta.size()) ||
      !write_file(argv[18], piv.data(), sizeof(int) * piv.size()))
    return 1;
  return status;
}
