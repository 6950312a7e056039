// Host validator behind check=true: fit error and FFT Coulomb error of an
// ISDF result, as LibISDF's checkcode computes them
// (src/common/checkcode.cpp:54-313, with src/common/fourier.cpp:26-198).
//
// Memory is kept small: the fit error streams over blocks of pair columns, so
// Z = psi_k psi_j is never stored.  The Coulomb check holds the K1*K2 pair
// spectra (as the reference does) and is meant for small grids.
// Work is split over std::thread workers; every reduction is done per
// worker and then summed in worker order, so results are deterministic for a
// fixed thread count.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/isdf_hip.h"
#include "isdf_internal.h"

namespace {

typedef std::complex<double> cplx;

int nthreads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  return (int)std::min(n, 32u);
}

template <class F>
void parallel_for(long long n, F f) {
  int nt = (int)std::min<long long>(nthreads(), std::max<long long>(1, n));
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++) {
    long long a = n * t / nt, b = n * (t + 1) / nt;
    th.emplace_back([=]() { f(t, a, b); });
  }
  for (auto& x : th) x.join();
}

// Forward 3-D DFT of real columns with FFTW's sign (e^{-i}), scaled by vol/Ng
// (fourier.cpp:191-196), full spectrum in grid order (what FFTR2C rebuilds by
// Hermitian symmetry, checkcode.cpp:9-53).  Separable direct DFT per axis.
struct Dft3 {
  int n[3];
  std::vector<cplx> tw[3];
  double scale;
  Dft3(const int ng[3], double vol) {
    for (int d = 0; d < 3; d++) {
      n[d] = ng[d];
      tw[d].resize(n[d]);
      for (int k = 0; k < n[d]; k++) {
        double a = -2.0 * M_PI * (double)k / (double)n[d];
        tw[d][k] = cplx(std::cos(a), std::sin(a));
      }
    }
    scale = vol / (double)((long long)n[0] * n[1] * n[2]);
  }
  // in: real column (Ng), out: complex spectrum (Ng); tmp: 2*Ng scratch.
  void run(const double* in, cplx* out, cplx* tmp) const {
    const int n0 = n[0], n1 = n[1], n2 = n[2];
    const long long ng = (long long)n0 * n1 * n2;
    cplx* a = tmp;
    // axis 0 (x, fastest)
    for (long long line = 0; line < (long long)n1 * n2; line++) {
      const double* src = in + line * n0;
      cplx* dst = a + line * n0;
      for (int k = 0; k < n0; k++) {
        cplx s(0.0, 0.0);
        for (int x = 0; x < n0; x++) s += src[x] * tw[0][(long long)k * x % n0];
        dst[k] = s;
      }
    }
    // axis 1 (y)
    std::vector<cplx> col(std::max(n1, n2));
    for (int z = 0; z < n2; z++)
      for (int x = 0; x < n0; x++) {
        for (int y = 0; y < n1; y++) col[y] = a[x + (long long)n0 * (y + (long long)n1 * z)];
        for (int k = 0; k < n1; k++) {
          cplx s(0.0, 0.0);
          for (int y = 0; y < n1; y++) s += col[y] * tw[1][(long long)k * y % n1];
          out[x + (long long)n0 * (k + (long long)n1 * z)] = s;
        }
      }
    // axis 2 (z)
    for (int y = 0; y < n1; y++)
      for (int x = 0; x < n0; x++) {
        for (int z = 0; z < n2; z++) col[z] = out[x + (long long)n0 * (y + (long long)n1 * z)];
        for (int k = 0; k < n2; k++) {
          cplx s(0.0, 0.0);
          for (int z = 0; z < n2; z++) s += col[z] * tw[2][(long long)k * z % n2];
          out[x + (long long)n0 * (y + (long long)n1 * k)] = s * scale;
        }
      }
    (void)ng;
  }
};

// |G|^2/2 on the FFT grid (fourier.cpp:55-94) and the truncated Coulomb kernel
// with radius 5 (checkcode.cpp:168-182).
std::vector<double> coulomb_kernel(const int ng[3], const double len[3]) {
  std::vector<double> K[3];
  for (int d = 0; d < 3; d++) {
    K[d].resize(ng[d]);
    for (int i = 0; i <= ng[d] / 2; i++) K[d][i] = i * 2.0 * M_PI / len[d];
    for (int i = ng[d] / 2 + 1; i < ng[d]; i++) K[d][i] = (i - ng[d]) * 2.0 * M_PI / len[d];
  }
  const double radius = 5.0;
  std::vector<double> v((size_t)ng[0] * ng[1] * ng[2]);
  size_t p = 0;
  for (int k = 0; k < ng[2]; k++)
    for (int j = 0; j < ng[1]; j++)
      for (int i = 0; i < ng[0]; i++) {
        double g = (K[0][i] * K[0][i] + K[1][j] * K[1][j] + K[2][k] * K[2][k]) / 2.0;
        v[p++] = (g < 1e-8) ? 4.0 * M_PI * radius * radius
                            : 8.0 * M_PI / (2.0 * g) * (1 - std::cos(std::sqrt(2.0 * g) * radius));
      }
  return v;
}

}  // namespace

extern "C" int isdf_hip_check(const isdf_problem* p, const double* psi, const double* theta,
                              const int* piv, int do_coulomb, isdf_check* out) {
  using isdf_hip::set_error;
  if (!p || !psi || !theta || !piv || !out) return set_error(ISDF_EINVAL, "NULL argument");
  const long long ng = (long long)p->ngrid[0] * p->ngrid[1] * p->ngrid[2];
  const long long ld = p->ld_psi ? p->ld_psi : ng;
  const int rk = isdf_hip_rank(p);
  const int s1 = p->nocc - p->nv1, K1 = p->nv1 + p->nc1;
  const int s2 = p->nocc - p->nv2, K2 = p->nv2 + p->nc2;
  if (rk <= 0 || K1 <= 0 || K2 <= 0 || s1 < 0 || s2 < 0)
    return set_error(ISDF_EINVAL, "invalid problem");
  const long long npair = (long long)K1 * K2;
  const int nt = nthreads();

  // Fit error, checkcode.cpp:62-166.
  {
    const int blk = 32;
    std::vector<double> cm((size_t)rk * blk);
    std::vector<double> e_t(nt), z_t(nt);
    double err = 0.0, nrm = 0.0;
    for (long long p0 = 0; p0 < npair; p0 += blk) {
      const int nb = (int)std::min<long long>(blk, npair - p0);
      for (int b = 0; b < nb; b++) {
        const long long q = p0 + b;
        const int k = (int)(q / K2), j = (int)(q % K2);
        for (int mu = 0; mu < rk; mu++)
          cm[(size_t)b * rk + mu] = psi[(long long)(s1 + k) * ld + piv[mu]] *
                                    psi[(long long)(s2 + j) * ld + piv[mu]];
      }
      std::fill(e_t.begin(), e_t.end(), 0.0);
      std::fill(z_t.begin(), z_t.end(), 0.0);
      parallel_for(ng, [&](int t, long long a, long long bnd) {
        double acc[32];
        double e = 0.0, z = 0.0;
        for (long long r = a; r < bnd; r++) {
          for (int b = 0; b < nb; b++) acc[b] = 0.0;
          for (int mu = 0; mu < rk; mu++) {
            const double tv = theta[(long long)mu * ng + r];
            const double* c = &cm[(size_t)mu];
            for (int b = 0; b < nb; b++) acc[b] += tv * c[(size_t)b * rk];
          }
          for (int b = 0; b < nb; b++) {
            const long long q = p0 + b;
            const int k = (int)(q / K2), j = (int)(q % K2);
            const double zz = psi[(long long)(s1 + k) * ld + r] * psi[(long long)(s2 + j) * ld + r];
            const double d = zz - acc[b];
            e += d * d;
            z += zz * zz;
          }
        }
        e_t[t] = e;
        z_t[t] = z;
      });
      for (int t = 0; t < nt; t++) {
        err += e_t[t];
        nrm += z_t[t];
      }
    }
    out->fit_error = std::sqrt(err) / std::sqrt(nrm);
    out->z_norm = std::sqrt(nrm);
  }
  out->coulomb_error = -1.0;
  out->coulomb_norm = -1.0;
  if (!do_coulomb) return ISDF_OK;

  // Coulomb error, checkcode.cpp:168-312.
  const double vol = p->length[0] * p->length[1] * p->length[2];
  Dft3 dft(p->ngrid, vol);
  std::vector<double> V = coulomb_kernel(p->ngrid, p->length);
  const long long ncol = npair + rk;
  std::vector<cplx> spec;
  try {
    spec.resize((size_t)(ncol * ng));
  } catch (...) {
    return set_error(ISDF_ENOMEM, "Coulomb check: spectra do not fit in host memory");
  }
  parallel_for(ncol, [&](int, long long a, long long b) {
    std::vector<double> col((size_t)ng);
    std::vector<cplx> tmp((size_t)ng);
    for (long long c = a; c < b; c++) {
      if (c < npair) {
        const int k = (int)(c / K2), j = (int)(c % K2);
        for (long long r = 0; r < ng; r++)
          col[r] = psi[(long long)(s1 + k) * ld + r] * psi[(long long)(s2 + j) * ld + r];
      } else {
        const double* t = theta + (c - npair) * ng;
        std::copy(t, t + ng, col.begin());
      }
      dft.run(col.data(), &spec[(size_t)(c * ng)], tmp.data());
    }
  });
  // Re(a^H diag(V) b) for all column pairs of a block.
  auto gram = [&](long long c0, long long n, std::vector<double>& G) {
    G.assign((size_t)(n * n), 0.0);
    parallel_for(n, [&](int, long long a, long long b) {
      for (long long i = a; i < b; i++) {
        const cplx* x = &spec[(size_t)((c0 + i) * ng)];
        for (long long j = 0; j < n; j++) {
          const cplx* y = &spec[(size_t)((c0 + j) * ng)];
          double s = 0.0;
          for (long long r = 0; r < ng; r++) s += V[r] * (x[r].real() * y[r].real() + x[r].imag() * y[r].imag());
          G[(size_t)(j * n + i)] = s;
        }
      }
    });
  };
  std::vector<double> ijVkl, M;
  gram(0, npair, ijVkl);
  gram(npair, rk, M);
  // C(mu, pair) = psi_k(r_mu) psi_j(r_mu) (checkcode.cpp:129-139); C^T M C.
  std::vector<double> C((size_t)rk * npair);
  for (long long q = 0; q < npair; q++) {
    const int k = (int)(q / K2), j = (int)(q % K2);
    for (int mu = 0; mu < rk; mu++)
      C[(size_t)(q * rk + mu)] =
          psi[(long long)(s1 + k) * ld + piv[mu]] * psi[(long long)(s2 + j) * ld + piv[mu]];
  }
  std::vector<double> MC((size_t)rk * npair);
  parallel_for(npair, [&](int, long long a, long long b) {
    for (long long q = a; q < b; q++)
      for (int i = 0; i < rk; i++) {
        double s = 0.0;
        for (int l = 0; l < rk; l++) s += M[(size_t)l * rk + i] * C[(size_t)(q * rk + l)];
        MC[(size_t)(q * rk + i)] = s;
      }
  });
  std::vector<double> e_t(nt), n_t(nt);
  parallel_for(npair, [&](int t, long long a, long long b) {
    double e = 0.0, nn = 0.0;
    for (long long q2 = a; q2 < b; q2++)
      for (long long q1 = 0; q1 < npair; q1++) {
        double s = 0.0;
        for (int i = 0; i < rk; i++) s += C[(size_t)(q1 * rk + i)] * MC[(size_t)(q2 * rk + i)];
        const double ref = ijVkl[(size_t)(q2 * npair + q1)];
        const double d = ref - s;
        e += d * d;
        nn += ref * ref;
      }
    e_t[t] = e;
    n_t[t] = nn;
  });
  double e = 0.0, nn = 0.0;
  for (int t = 0; t < nt; t++) {
    e += e_t[t];
    nn += n_t[t];
  }
  out->coulomb_error = std::sqrt(e) / std::sqrt(nn);
  out->coulomb_norm = std::sqrt(nn);
  return ISDF_OK;
}
