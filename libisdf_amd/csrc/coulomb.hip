// Coulomb check of an ISDF result on the device: the second half of LibISDF's
// check=true validator (src/common/checkcode.cpp:168-312).
//
//   ijVkl(a, b) = Re sum_G conj(rho_a(G)) V(G) rho_b(G),  rho_a = psi_k psi_j (pair a = k*K2 + j)
//   M(mu, nu)   = Re sum_G conj(zeta_mu(G)) V(G) zeta_nu(G), zeta_mu = column mu of Theta
//   error       = ||ijVkl - C^T M C||_F / ||ijVkl||_F,  C(mu, a) = rho_a(r_mu)
//
// with rho(G) the forward 3-D DFT (sign e^{-i}, scaled by vol/Ng, the full
// spectrum as FFTR2C rebuilds it, checkcode.cpp:9-53 and fourier.cpp:191-196)
// and V the truncated Coulomb kernel of radius 5 (checkcode.cpp:168-182).
//
// Layout: every spectrum, times sqrt(V(G)), is stored g-major as two real rows
// per grid point, Rt[g][c] = sqrt(V) Re, Rt[Ng + g][c] = sqrt(V) Im, so that
// both Gram matrices are plain FP64 MFMA products Rt^T Rt over K = 2 Ng (split
// over K into slices summed in a fixed order).  The DFT is separable: one pass
// per axis, a mixed-radix Stockham FFT of whole lines in LDS (k_fft_axis;
// an axis whose length has a prime factor above 13, or whose lines would need
// more than 64 KB of LDS, falls back to a direct sum over its line,
// k_dft_axis), with the host validator's twiddle table.
// Every sum has a fixed order, so the result is deterministic.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/isdf_hip.h"
#include "isdf_internal.h"
#include "mfma_gemm.h"

namespace isdf_hip {

int problem_rank(const isdf_problem* p);
int validate_problem(const isdf_problem* p, long long* ng_out);
std::mutex& solve_mutex();
int fit_error_unlocked(const isdf_problem* p, const double* d_psi, const double* d_theta,
                       const int* h_piv, hipStream_t st, isdf_check* out);

namespace {

__host__ __device__ inline long long cdivl(long long a, long long b) { return (a + b - 1) / b; }

constexpr double kPi = 3.14159265358979323846;

// sqrt(V(G)) on the FFT grid: |G|^2/2 as fourier.cpp:55-94 forms it, V as
// checkcode.cpp:168-182 (radius 5).
__global__ __launch_bounds__(256) void k_coul_sqrtv(int n0, int n1, int n2, double L0, double L1,
                                                    double L2, double* __restrict__ sv) {
  const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long ng = (long long)n0 * n1 * n2;
  if (g >= ng) return;
  const int i = (int)(g % n0), j = (int)((g / n0) % n1), k = (int)(g / ((long long)n0 * n1));
  const double kx = (i <= n0 / 2 ? i : i - n0) * 2.0 * kPi / L0;
  const double ky = (j <= n1 / 2 ? j : j - n1) * 2.0 * kPi / L1;
  const double kz = (k <= n2 / 2 ? k : k - n2) * 2.0 * kPi / L2;
  const double gkk = (kx * kx + ky * ky + kz * kz) / 2.0;
  const double radius = 5.0;
  const double v = gkk < 1e-8 ? 4.0 * kPi * radius * radius
                              : 8.0 * kPi / (2.0 * gkk) * (1 - cos(sqrt(2.0 * gkk) * radius));
  sv[g] = sqrt(v);
}

// Columns c in [c0, c0 + nb) as complex data (imaginary part 0): the pair
// densities for c < npair, Theta's columns after them.
__global__ __launch_bounds__(256) void k_coul_cols(const double* __restrict__ psi, long long ld,
                                                   int s1, int s2, int K2, long long npair,
                                                   const double* __restrict__ theta,
                                                   long long ng, long long c0, int nb,
                                                   double2* __restrict__ X) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)nb * ng) return;
  const long long c = c0 + e / ng, r = e % ng;
  double v;
  if (c < npair) {
    const int k = (int)(c / K2), j = (int)(c % K2);
    v = psi[(long long)(s1 + k) * ld + r] * psi[(long long)(s2 + j) * ld + r];
  } else {
    v = theta[(c - npair) * ng + r];
  }
  X[e] = make_double2(v, 0.0);
}

// One axis of the DFT for nb columns: out(.., k, ..) = sum_t in(.., t, ..) tw[(k t) mod n],
// t ascending.  LAST: the z pass, which also applies vol/Ng and sqrt(V) and
// writes the g-major real rows of Rt (columns c0.. of ncol).
template <int AXIS, bool LAST>
__global__ __launch_bounds__(256) void k_dft_axis(const double2* __restrict__ in,
                                                  double2* __restrict__ out, int n0, int n1,
                                                  int n2, int nb, const double2* __restrict__ tw,
                                                  double scale, const double* __restrict__ sv,
                                                  double* __restrict__ Rt, long long ncol,
                                                  long long c0) {
  const long long ng = (long long)n0 * n1 * n2;
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)nb * ng) return;
  const long long col = e / ng, g = e % ng;
  const int x = (int)(g % n0), y = (int)((g / n0) % n1), z = (int)(g / ((long long)n0 * n1));
  int n, k;
  long long stride, base;
  if constexpr (AXIS == 0) {
    n = n0, k = x, stride = 1, base = g - x;
  } else if constexpr (AXIS == 1) {
    n = n1, k = y, stride = n0, base = g - (long long)y * n0;
  } else {
    n = n2, k = z, stride = (long long)n0 * n1, base = g - (long long)z * n0 * n1;
  }
  const double2* src = in + col * ng + base;
  double re = 0.0, im = 0.0;
  int kt = 0;
  for (int t = 0; t < n; t++) {
    const double2 a = src[t * stride];
    const double2 w = tw[kt];
    re += a.x * w.x - a.y * w.y;
    im += a.x * w.y + a.y * w.x;
    kt += k;
    if (kt >= n) kt -= n;
  }
  if constexpr (LAST) {
    const double f = scale * sv[g];
    Rt[g * ncol + c0 + col] = re * f;
    Rt[(ng + g) * ncol + c0 + col] = im * f;
  } else {
    out[e] = make_double2(re, im);
  }
}

// Radices of a mixed-radix FFT of length n: fours, a two, then odd primes up to
// kMaxRadix (n with a larger prime factor: count = -1, the direct DFT above
// serves it).
constexpr int kMaxRadix = 13, kMaxPass = 16;
// LDS of one FFT workgroup at most (the default dynamic LDS limit); a longer
// axis takes the direct DFT.
constexpr size_t kFftMaxLds = 65536;
struct FftPlan {
  int count;
  int radix[kMaxPass];
};

// One Stockham pass of radix R over TL lines of length n held in LDS rows of
// ld entries (A -> B); Ns is the product of the earlier radices.  Butterfly j
// of a line takes a[j + r n/R] times w^{(j mod Ns) r n/(Ns R)} (w = T[1] =
// e^{-2 pi i/n}), transforms the R values (radix 2 and 4 by their closed
// forms, odd R as a direct length-R DFT with w_R = w^{n/R}) and writes them to
// b[(j / Ns) Ns R + j mod Ns + r Ns].  R is a template argument, so the R
// values stay in registers.
template <int R>
__device__ __forceinline__ void fft_pass(const double2* A, double2* B, const double2* T, int TL,
                                         int ld, int n, int Ns) {
  const int nR = n / R, step = n / (Ns * R);
  for (int w = threadIdx.x; w < TL * nR; w += 256) {
    const int li = w / nR, j = w - li * nR, jj = j % Ns;
    const double2* a = A + li * ld;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const double2 x = a[j + r * nR];
      const double2 c = T[(jj * r * step) % n];
      v[r] = make_double2(x.x * c.x - x.y * c.y, x.x * c.y + x.y * c.x);
    }
    double2 o[R];
    if constexpr (R == 2) {
      o[0] = make_double2(v[0].x + v[1].x, v[0].y + v[1].y);
      o[1] = make_double2(v[0].x - v[1].x, v[0].y - v[1].y);
    } else if constexpr (R == 4) {  // w_4 = -i
      const double2 s0 = make_double2(v[0].x + v[2].x, v[0].y + v[2].y);
      const double2 d0 = make_double2(v[0].x - v[2].x, v[0].y - v[2].y);
      const double2 s1 = make_double2(v[1].x + v[3].x, v[1].y + v[3].y);
      const double2 d1 = make_double2(v[1].y - v[3].y, v[3].x - v[1].x);  // -i (v1 - v3)
      o[0] = make_double2(s0.x + s1.x, s0.y + s1.y);
      o[2] = make_double2(s0.x - s1.x, s0.y - s1.y);
      o[1] = make_double2(d0.x + d1.x, d0.y + d1.y);
      o[3] = make_double2(d0.x - d1.x, d0.y - d1.y);
    } else {
#pragma unroll
      for (int k = 0; k < R; k++) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int r = 0; r < R; r++) {
          const double2 c = T[((k * r) % R) * nR];
          re += v[r].x * c.x - v[r].y * c.y;
          im += v[r].x * c.y + v[r].y * c.x;
        }
        o[k] = make_double2(re, im);
      }
    }
    double2* b = B + li * ld + (j / Ns) * Ns * R + jj;
#pragma unroll
    for (int r = 0; r < R; r++) b[r * Ns] = o[r];
  }
}

// One axis of the DFT for nb columns as a Stockham autosort FFT in LDS (the
// same transform as k_dft_axis: out_k = sum_t in_t e^{-2 pi i k t / n}).  A
// workgroup loads TL whole lines (element t of line l at base(l) + t stride),
// runs the passes between two LDS buffers (pass p: radix R, Ns the product of
// the earlier radices; butterfly j takes in[j + r n/R] times the twiddle
// w^{(j mod Ns) r n/(Ns R)}, transforms them, and writes out[(j / Ns) Ns R +
// j mod Ns + r Ns]) and writes the lines back; LAST as in k_dft_axis.  The
// twiddles w^q = tw[q] are the host validator's table.
template <int AXIS, bool LAST>
__global__ __launch_bounds__(256) void k_fft_axis(const double2* __restrict__ in,
                                                  double2* __restrict__ out, int n0, int n1,
                                                  int n2, int nb, const double2* __restrict__ tw,
                                                  FftPlan plan, int TL, double scale,
                                                  const double* __restrict__ sv,
                                                  double* __restrict__ Rt, long long ncol,
                                                  long long c0) {
  extern __shared__ double2 sm[];
  const long long ng = (long long)n0 * n1 * n2;
  const int n = AXIS == 0 ? n0 : (AXIS == 1 ? n1 : n2);
  const int ld = n + 1;  // LDS row of a line, padded
  double2* A = sm;
  double2* B = sm + (size_t)TL * ld;
  double2* T = sm + 2 * (size_t)TL * ld;
  const long long lines_per_col = ng / n;
  const long long nlines = lines_per_col * nb;
  const long long l0 = (long long)blockIdx.x * TL;
  // Line l (x fastest among the other two axes, then the column): its
  // first element and the stride along the axis.
  auto line_at = [&](long long l, long long* base, long long* stride, long long* col) {
    const long long c = l / lines_per_col, r = l % lines_per_col;
    *col = c;
    if constexpr (AXIS == 0) {
      *base = c * ng + r * n0;
      *stride = 1;
    } else if constexpr (AXIS == 1) {
      const long long x = r % n0, z = r / n0;
      *base = c * ng + z * (long long)n0 * n1 + x;
      *stride = n0;
    } else {
      *base = c * ng + r;  // r = x + n0 y
      *stride = (long long)n0 * n1;
    }
  };
  // Element e of the tile: along the line fastest for axis 0 (contiguous),
  // across lines fastest otherwise (adjacent lines are adjacent in memory).
  auto tile_elem = [&](int e, int* li, int* t) {
    if constexpr (AXIS == 0) {
      *li = e / n;
      *t = e - *li * n;
    } else {
      *t = e / TL;
      *li = e - *t * TL;
    }
  };
  for (int q = threadIdx.x; q < n; q += 256) T[q] = tw[q];
  for (int e = threadIdx.x; e < TL * n; e += 256) {
    int li, t;
    tile_elem(e, &li, &t);
    const long long l = l0 + li;
    if (l >= nlines) continue;
    long long base, stride, col;
    line_at(l, &base, &stride, &col);
    A[li * ld + t] = in[base + t * stride];
  }
  __syncthreads();
  int Ns = 1;
  for (int p = 0; p < plan.count; p++) {
    const int R = plan.radix[p];
    switch (R) {
      case 2: fft_pass<2>(A, B, T, TL, ld, n, Ns); break;
      case 3: fft_pass<3>(A, B, T, TL, ld, n, Ns); break;
      case 4: fft_pass<4>(A, B, T, TL, ld, n, Ns); break;
      case 5: fft_pass<5>(A, B, T, TL, ld, n, Ns); break;
      case 7: fft_pass<7>(A, B, T, TL, ld, n, Ns); break;
      case 11: fft_pass<11>(A, B, T, TL, ld, n, Ns); break;
      default: fft_pass<13>(A, B, T, TL, ld, n, Ns); break;
    }
    __syncthreads();
    double2* s = A;
    A = B;
    B = s;
    Ns *= R;
  }
  for (int e = threadIdx.x; e < TL * n; e += 256) {
    int li, t;
    tile_elem(e, &li, &t);
    const long long l = l0 + li;
    if (l >= nlines) continue;
    long long base, stride, col;
    line_at(l, &base, &stride, &col);
    const long long idx = base + t * stride;
    const double2 v = A[li * ld + t];
    if constexpr (LAST) {
      const long long g = idx - col * ng;
      const double f = scale * sv[g];
      Rt[g * ncol + c0 + col] = v.x * f;
      Rt[(ng + g) * ncol + c0 + col] = v.y * f;
    } else {
      out[idx] = v;
    }
  }
}

FftPlan fft_plan(int n) {
  FftPlan pl{};
  int m = n;
  auto add = [&](int r) {
    if (pl.count >= kMaxPass) return false;
    pl.radix[pl.count++] = r;
    return true;
  };
  while (m % 4 == 0) {
    add(4);
    m /= 4;
  }
  while (m % 2 == 0) {
    add(2);
    m /= 2;
  }
  for (int p = 3; m > 1 && p <= kMaxRadix; p += 2) {
    while (m % p == 0) {
      add(p);
      m /= p;
    }
  }
  if (m > 1) pl.count = -1;
  return pl;
}

// Lines per FFT workgroup: the two tiles and the twiddles within ~48 KB of LDS.
int fft_lines(int n) { return std::max(1, std::min(128, 1536 / (n + 1))); }

// C(mu, a) = rho_a(r_mu), stored mu-major: Cm[mu * npair + a].
__global__ __launch_bounds__(256) void k_coul_cmat(const double* __restrict__ psi, long long ld,
                                                   const int* __restrict__ piv, int rk, int s1,
                                                   int s2, int K2, long long npair,
                                                   double* __restrict__ Cm) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)rk * npair) return;
  const int mu = (int)(e / npair);
  const long long a = e % npair;
  const int k = (int)(a / K2), j = (int)(a % K2);
  const long long g = piv[mu];
  Cm[e] = psi[(long long)(s1 + k) * ld + g] * psi[(long long)(s2 + j) * ld + g];
}

// out[s][n][m] (column-major M x N per slice s = blockIdx.z) = sum over the
// slice's k range of A(m, k) B(k, n); A(m, k) at A[m + k lda], B as the tile
// type reads it (B_KMAJOR: B[k ldb + n], else B[k + n ldb]).
template <class Tile>
__global__ __launch_bounds__(Tile::NT) void k_gemm_slices(const double* __restrict__ A,
                                                          long long lda,
                                                          const double* __restrict__ B,
                                                          long long ldb, int M, int N, long long K,
                                                          long long kslice,
                                                          double* __restrict__ out) {
  __shared__ double lds[Tile::LDS_DOUBLES];
  const int m0 = blockIdx.x * Tile::BM, n0 = blockIdx.y * Tile::BN;
  const long long k0 = (long long)blockIdx.z * kslice;
  const int klen = (int)std::min<long long>(kslice, K - k0);
  const int mlim = std::min(Tile::BM, M - m0), nlim = std::min(Tile::BN, N - n0);
  typename Tile::Acc acc;
  Tile::zero(acc);
  const double* Ab = A + k0 * lda + m0;
  const double* Bb = Tile::B_KMAJOR ? B + k0 * ldb + n0 : B + (long long)n0 * ldb + k0;
  Tile::run(Ab, lda, Bb, ldb, klen, mlim, nlim, lds, acc);
  double* o = out + (long long)blockIdx.z * M * N;
  Tile::epilogue(acc, [&](int m, int n, double v) {
    if (m < mlim && n < nlim) o[(long long)(n0 + n) * M + m0 + m] = v;
  });
}
typedef MfmaTile<64, 64, 16, 2, 2, true> CoulTileK;   // B k-major
typedef MfmaTile<64, 64, 16, 2, 2, false> CoulTileN;  // B k-contiguous

// dst[i] = sum over s = 0..S-1, in order, of parts[s * count + i].
__global__ __launch_bounds__(256) void k_sum_slices(const double* __restrict__ parts, int S,
                                                    long long count, double* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  double s = 0.0;
  for (int q = 0; q < S; q++) s += parts[(long long)q * count + i];
  dst[i] = s;
}

// part[2b] = sum of (G - A)^2 and part[2b+1] = sum of G^2 over workgroup b's
// 256 * 16 entries (fixed order: per thread ascending, then a fixed tree).
__global__ __launch_bounds__(256) void k_coul_err(const double* __restrict__ G,
                                                  const double* __restrict__ A, long long count,
                                                  double* __restrict__ part) {
  __shared__ double se[256], sn[256];
  double e = 0.0, n = 0.0;
  const long long base = (long long)blockIdx.x * 256 * 16;
  for (int q = 0; q < 16; q++) {
    const long long i = base + (long long)q * 256 + threadIdx.x;
    if (i < count) {
      const double d = G[i] - A[i];
      e += d * d;
      n += G[i] * G[i];
    }
  }
  se[threadIdx.x] = e;
  sn[threadIdx.x] = n;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) {
      se[threadIdx.x] += se[threadIdx.x + h];
      sn[threadIdx.x] += sn[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * (long long)blockIdx.x] = se[0];
    part[2 * (long long)blockIdx.x + 1] = sn[0];
  }
}

// One workgroup: out[0..1] = the two sums of part over nb slots (fixed order).
__global__ __launch_bounds__(1024) void k_pair_total(const double* __restrict__ part, int nb,
                                                     double* __restrict__ out) {
  __shared__ double se[1024], sn[1024];
  double e = 0.0, n = 0.0;
  for (int b = threadIdx.x; b < nb; b += 1024) {
    e += part[2 * (long long)b];
    n += part[2 * (long long)b + 1];
  }
  se[threadIdx.x] = e;
  sn[threadIdx.x] = n;
  __syncthreads();
  for (int h = 512; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) {
      se[threadIdx.x] += se[threadIdx.x + h];
      sn[threadIdx.x] += sn[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = se[0];
    out[1] = sn[0];
  }
}

// Grow-only device buffer of the Coulomb check, one per device.
struct CoulBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};
std::vector<CoulBuf> g_coul;

int coul_buffer(size_t bytes, void** out) {
  int dev = 0;
  ISDF_HIP_TRY(hipGetDevice(&dev));
  if ((int)g_coul.size() <= dev) g_coul.resize(dev + 1);
  CoulBuf& b = g_coul[dev];
  if (b.bytes < bytes) {
    if (b.ptr) ISDF_HIP_TRY(hipFree(b.ptr));
    b.ptr = nullptr;
    b.bytes = 0;
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      b.ptr = nullptr;
      return set_error(ISDF_ENOMEM, "Coulomb check: device allocation of " +
                                        std::to_string(bytes) +
                                        " bytes failed: " + hipGetErrorString(e));
    }
    b.bytes = bytes;
  }
  *out = b.ptr;
  return 0;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Slices of K for an M x N Gram: enough workgroups to cover the chip
// (~2048), each slice a multiple of the tile's K chunk.
long long gram_slices(int M, int N, long long K, long long* kslice) {
  const long long tiles = cdivl(M, 64) * cdivl(N, 64);
  long long S = std::max<long long>(1, std::min<long long>(cdivl(2048, tiles), K / 512));
  long long ks = cdivl(cdivl(K, S), 16) * 16;
  S = cdivl(K, ks);
  *kslice = ks;
  return S;
}

template <class Tile>
int launch_gemm(const double* A, long long lda, const double* B, long long ldb, int M, int N,
                long long K, long long kslice, long long S, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_gemm_slices<Tile>, dim3((unsigned)cdivl(M, 64), (unsigned)cdivl(N, 64),
                                               (unsigned)S),
                     dim3(Tile::NT), 0, st, A, lda, B, ldb, M, N, K, kslice, out);
  ISDF_HIP_LAUNCHED("k_gemm_slices");
  return 0;
}

int coulomb_unlocked(const isdf_problem* p, const double* d_psi, const double* d_theta,
                     const int* h_piv, hipStream_t st, isdf_check* out) {
  long long ng = 0;
  if (int rc = validate_problem(p, &ng)) return rc;
  const int rk = problem_rank(p);
  const long long ld = p->ld_psi ? p->ld_psi : ng;
  const int s1 = p->nocc - p->nv1, K1 = p->nv1 + p->nc1;
  const int s2 = p->nocc - p->nv2, K2 = p->nv2 + p->nc2;
  const long long npair = (long long)K1 * K2;
  const long long ncol = npair + rk;
  if (npair > 65536) return set_error(ISDF_EINVAL, "Coulomb check: more than 65536 pairs");
  const int n0 = p->ngrid[0], n1 = p->ngrid[1], n2 = p->ngrid[2];

  long long ks_g = 0, ks_m = 0;
  const long long S_g = gram_slices((int)npair, (int)npair, 2 * ng, &ks_g);
  const long long S_m = gram_slices(rk, rk, 2 * ng, &ks_m);
  // Columns per DFT block: the two complex work buffers stay within 1 GiB.
  const long long nb = std::max<long long>(1, std::min<long long>(ncol, (1LL << 30) / (32 * ng)));
  const long long nerr = cdivl(npair * npair, 256 * 16);

  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  };
  const size_t o_sv = take(sizeof(double) * ng);
  const size_t o_tw = take(sizeof(double2) * (size_t)(n0 + n1 + n2));
  const size_t o_X = take(sizeof(double2) * (size_t)(nb * ng));
  const size_t o_Y = take(sizeof(double2) * (size_t)(nb * ng));
  const size_t o_Rt = take(sizeof(double) * (size_t)(2 * ng * ncol));
  const size_t o_piv = take(sizeof(int) * rk);
  const size_t o_Cm = take(sizeof(double) * (size_t)(rk * npair));
  // With one K slice the slice buffer is the result itself.
  const size_t o_Ms = take(sizeof(double) * (size_t)(S_m * rk * rk));
  const size_t o_M = S_m > 1 ? take(sizeof(double) * (size_t)rk * rk) : o_Ms;
  const size_t o_W = take(sizeof(double) * (size_t)(rk * npair));
  const size_t o_A = take(sizeof(double) * (size_t)(npair * npair));
  const size_t o_Gs = take(sizeof(double) * (size_t)(S_g * npair * npair));
  const size_t o_G = S_g > 1 ? take(sizeof(double) * (size_t)(npair * npair)) : o_Gs;
  const size_t o_part = take(2 * sizeof(double) * (size_t)nerr);
  const size_t o_out = take(2 * sizeof(double));
  void* base = nullptr;
  if (int rc = coul_buffer(off, &base)) return rc;
  char* b = (char*)base;
  double* sv = (double*)(b + o_sv);
  double2* tw = (double2*)(b + o_tw);
  double2* X = (double2*)(b + o_X);
  double2* Y = (double2*)(b + o_Y);
  double* Rt = (double*)(b + o_Rt);
  int* d_piv = (int*)(b + o_piv);
  double* Cm = (double*)(b + o_Cm);
  double* Ms = (double*)(b + o_Ms);
  double* Mm = (double*)(b + o_M);
  double* W = (double*)(b + o_W);
  double* Am = (double*)(b + o_A);
  double* Gs = (double*)(b + o_Gs);
  double* G = (double*)(b + o_G);
  double* part = (double*)(b + o_part);
  double* d_out = (double*)(b + o_out);

  // Twiddles e^{-2 pi i k / n} per axis, as the host validator forms them.
  std::vector<double2> htw((size_t)(n0 + n1 + n2));
  {
    size_t q = 0;
    for (int d = 0; d < 3; d++) {
      const int n = p->ngrid[d];
      for (int k = 0; k < n; k++) {
        const double a = -2.0 * kPi * (double)k / (double)n;
        htw[q++] = make_double2(std::cos(a), std::sin(a));
      }
    }
  }
  ISDF_HIP_TRY(hipMemcpyAsync(tw, htw.data(), sizeof(double2) * htw.size(),
                              hipMemcpyHostToDevice, st));
  ISDF_HIP_TRY(hipMemcpyAsync(d_piv, h_piv, sizeof(int) * rk, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_coul_sqrtv, dim3((unsigned)cdivl(ng, 256)), dim3(256), 0, st, n0, n1, n2,
                     p->length[0], p->length[1], p->length[2], sv);
  ISDF_HIP_LAUNCHED("k_coul_sqrtv");

  // Spectra of every column into Rt.
  const FftPlan plan[3] = {fft_plan(n0), fft_plan(n1), fft_plan(n2)};
  const double scale = p->length[0] * p->length[1] * p->length[2] / (double)ng;
  for (long long c0 = 0; c0 < ncol; c0 += nb) {
    const int w = (int)std::min<long long>(nb, ncol - c0);
    const dim3 grid((unsigned)cdivl((long long)w * ng, 256));
    hipLaunchKernelGGL(k_coul_cols, grid, dim3(256), 0, st, d_psi, ld, s1, s2, K2, npair, d_theta,
                       ng, c0, w, X);
    // Per axis: the FFT when its length factors into radices up to
    // kMaxRadix and its lines fit the LDS bound, else the direct DFT.
    auto axis = [&](auto ax, auto last, const double2* src, double2* dst, const double2* t) {
      constexpr int AX = decltype(ax)::value;
      constexpr bool LAST = decltype(last)::value;
      const int n = p->ngrid[AX];
      const int TL = fft_lines(n);
      const size_t lds = sizeof(double2) * (2 * (size_t)TL * (n + 1) + n);
      if (plan[AX].count >= 0 && lds <= kFftMaxLds) {
        const long long nlines = (long long)w * (ng / n);
        hipLaunchKernelGGL((k_fft_axis<AX, LAST>), dim3((unsigned)cdivl(nlines, TL)), dim3(256),
                           lds, st, src, dst, n0, n1, n2, w, t, plan[AX], TL, scale,
                           (const double*)sv, Rt, ncol, c0);
      } else {
        hipLaunchKernelGGL((k_dft_axis<AX, LAST>), grid, dim3(256), 0, st, src, dst, n0, n1, n2, w,
                           t, scale, (const double*)sv, Rt, ncol, c0);
      }
    };
    axis(std::integral_constant<int, 0>(), std::false_type(), (const double2*)X, Y,
         (const double2*)tw);
    axis(std::integral_constant<int, 1>(), std::false_type(), (const double2*)Y, X,
         (const double2*)(tw + n0));
    axis(std::integral_constant<int, 2>(), std::true_type(), (const double2*)X, Y,
         (const double2*)(tw + n0 + n1));
    ISDF_HIP_LAUNCHED("k_fft_axis");
  }

  // M = Theta spectra Gram (rk x rk); W = M C (b-major); A = C^T W.
  if (int rc = launch_gemm<CoulTileK>(Rt + npair, ncol, Rt + npair, ncol, rk, rk, 2 * ng, ks_m,
                                      S_m, Ms, st))
    return rc;
  if (S_m > 1) {
    hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)cdivl((long long)rk * rk, 256)), dim3(256),
                       0, st, (const double*)Ms, (int)S_m, (long long)rk * rk, Mm);
    ISDF_HIP_LAUNCHED("k_sum_slices");
  }
  hipLaunchKernelGGL(k_coul_cmat, dim3((unsigned)cdivl((long long)rk * npair, 256)), dim3(256), 0,
                     st, d_psi, ld, (const int*)d_piv, rk, s1, s2, K2, npair, Cm);
  ISDF_HIP_LAUNCHED("k_coul_cmat");
  if (int rc = launch_gemm<CoulTileK>(Mm, rk, Cm, npair, rk, (int)npair, rk, rk, 1, W, st))
    return rc;
  if (int rc = launch_gemm<CoulTileN>(Cm, npair, W, rk, (int)npair, (int)npair, rk, rk, 1, Am, st))
    return rc;
  // ijVkl = pair spectra Gram (npair x npair).
  if (int rc = launch_gemm<CoulTileK>(Rt, ncol, Rt, ncol, (int)npair, (int)npair, 2 * ng, ks_g,
                                      S_g, Gs, st))
    return rc;
  if (S_g > 1) {
    hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)cdivl(npair * npair, 256)), dim3(256), 0, st,
                       (const double*)Gs, (int)S_g, npair * npair, G);
    ISDF_HIP_LAUNCHED("k_sum_slices");
  }
  hipLaunchKernelGGL(k_coul_err, dim3((unsigned)nerr), dim3(256), 0, st, (const double*)G,
                     (const double*)Am, npair * npair, part);
  ISDF_HIP_LAUNCHED("k_coul_err");
  hipLaunchKernelGGL(k_pair_total, dim3(1), dim3(1024), 0, st, (const double*)part, (int)nerr,
                     d_out);
  ISDF_HIP_LAUNCHED("k_pair_total");
  double h[2] = {0.0, 0.0};
  ISDF_HIP_TRY(hipMemcpyAsync(h, d_out, sizeof(h), hipMemcpyDeviceToHost, st));
  ISDF_HIP_TRY(hipStreamSynchronize(st));
  out->coulomb_error = std::sqrt(h[0]) / std::sqrt(h[1]);
  out->coulomb_norm = std::sqrt(h[1]);
  return ISDF_OK;
}

}  // namespace

int release_coulomb_buffer() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return ISDF_EHIP;
  if ((int)g_coul.size() > dev && g_coul[dev].ptr) {
    if (hipFree(g_coul[dev].ptr) != hipSuccess) return ISDF_EHIP;
    g_coul[dev].ptr = nullptr;
    g_coul[dev].bytes = 0;
  }
  return ISDF_OK;
}

}  // namespace isdf_hip

extern "C" int isdf_hip_check_device(const isdf_problem* prob, const double* d_psi,
                                     const double* d_theta, const int* h_piv, int do_coulomb,
                                     void* stream, isdf_check* out) {
  std::lock_guard<std::mutex> lk(isdf_hip::solve_mutex());
  isdf_hip::set_error(0, "");
  int rc = isdf_hip::fit_error_unlocked(prob, d_psi, d_theta, h_piv, (hipStream_t)stream, out);
  if (rc == ISDF_OK && do_coulomb)
    rc = isdf_hip::coulomb_unlocked(prob, d_psi, d_theta, h_piv, (hipStream_t)stream, out);
  if (rc == ISDF_EHIP) (void)hipStreamSynchronize((hipStream_t)stream);
  return rc;
}
