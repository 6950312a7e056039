"""ISDF oracle: CPU restatement of LibISDF's CPU ISDF (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only as the checker.  The product path
(``libisdf_amd``) never imports, links or runs anything here.

``liboracle_isdf.so`` is the C++ restatement in ``isdf_oracle.cpp`` (built by
``make -C oracle``).  The FFT Coulomb check of ``src/common/checkcode.cpp:168-312``
is restated here in numpy, since it needs an FFT (the reference uses FFTW).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ISDF_ORACLE_LIB selects another build (the sanitizer build, oracle/_san/).
_LIB_PATH = os.environ.get("ISDF_ORACLE_LIB") or os.path.join(_HERE, "liboracle_isdf.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the C++ oracle (g++, -ffp-contract=off)."""
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_irank.restype = ctypes.c_int
        L.orc_irank.argtypes = [ctypes.c_int] * 5
        L.orc_weights.restype = None
        L.orc_weights.argtypes = [_dp, ctypes.c_long, ctypes.c_int] + [ctypes.c_int] * 5 + [_dp]
        L.orc_kmeans.restype = ctypes.c_int
        L.orc_kmeans.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                 ctypes.c_double, _dp, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _ip]
        L.orc_gather.restype = None
        L.orc_gather.argtypes = [_dp, ctypes.c_long, ctypes.c_int, _ip, ctypes.c_int, _dp]
        L.orc_zct_rows.restype = None
        L.orc_zct_rows.argtypes = [_dp, ctypes.c_long, ctypes.c_int, _dp, ctypes.c_int] + \
            [ctypes.c_int] * 5 + [ctypes.c_long, ctypes.c_long, _dp, ctypes.c_long]
        L.orc_zct_pick.restype = None
        L.orc_zct_pick.argtypes = [_dp, ctypes.c_long, ctypes.c_int, _dp, ctypes.c_int] + \
            [ctypes.c_int] * 5 + [_ip, ctypes.c_long, _dp, ctypes.c_long]
        L.orc_cct.restype = ctypes.c_double
        L.orc_cct.argtypes = [_dp, ctypes.c_long, _ip, ctypes.c_int, _dp]
        L.orc_potrf.restype = ctypes.c_int
        L.orc_potrf.argtypes = [_dp, ctypes.c_int]
        L.orc_potrs_rows.restype = None
        L.orc_potrs_rows.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_long, ctypes.c_long]
        L.orc_fit_error.restype = ctypes.c_double
        L.orc_fit_error.argtypes = [_dp, ctypes.c_long, ctypes.c_int, ctypes.c_int] + \
            [ctypes.c_int] * 5 + [_dp, ctypes.c_long, ctypes.c_int, _ip, _dp]
        L.orc_isdf.restype = ctypes.c_int
        L.orc_isdf.argtypes = [_dp, ctypes.c_long, _dp, _ip] + [ctypes.c_int] * 7 + \
            [_dp, _ip, ctypes.c_int, _ip, _ip, _ip]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_glibc_rand.restype = None
        L.orc_glibc_rand.argtypes = [ctypes.c_uint, ctypes.c_int, _ip]
        L.orc_random_shuffle.restype = None
        L.orc_random_shuffle.argtypes = [_ip, ctypes.c_long, ctypes.c_uint]
        L.orc_set_scalar.restype = None
        L.orc_set_scalar.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _psi_arr(psi):
    """psi as (nstate, Ng) C-contiguous float64, i.e. column-major Ng x nstate."""
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    assert psi.ndim == 2
    return psi


def irank(mu_points, nv1, nc1, nv2, nc2):
    return lib().orc_irank(mu_points, nv1, nc1, nv2, nc2)


def weights(psi, nocc, nv1, nc1, nv2, nc2):
    psi = _psi_arr(psi)
    ng = psi.shape[1]
    w = np.empty(ng, np.float64)
    lib().orc_weights(_d(psi), ng, ng, nocc, nv1, nc1, nv2, nc2, _d(w))
    return w


def kmeans(weight, rk, max_iter, length, ngrid, kmeans_tol=1e-5, df_tol=1e-8):
    """KMEAN at one rank.  Returns dict(piv, rk_kmeans, iters, trace, labels,
    centroids, npt, branch)."""
    weight = np.ascontiguousarray(weight, np.float64)
    n = weight.size
    piv = np.zeros(n, np.int32)
    rk_out = np.zeros(1, np.int32)
    iters = np.zeros(1, np.int32)
    trace = np.zeros(max(max_iter, 1), np.int32)
    labels = np.zeros(n, np.int32)
    cent = np.zeros(3 * rk, np.float64)
    npt = np.zeros(1, np.int32)
    L = np.asarray(length, np.float64)
    G = np.asarray(ngrid, np.int32)
    br = lib().orc_kmeans(n, _d(weight), rk, kmeans_tol, max_iter, df_tol, _d(L), _i(G),
                          _i(piv), _i(rk_out), _i(iters), _i(trace), _i(labels), _d(cent),
                          _i(npt))
    it = int(iters[0])
    return dict(piv=piv, rk_kmeans=int(rk_out[0]), iters=it, trace=trace[:it].copy(),
                labels=labels, centroids=cent, npt=int(npt[0]), branch=br)


def glibc_rand(seed, n):
    """The first n values of glibc rand() after srand(seed), as the oracle restates it."""
    out = np.zeros(n, np.int32)
    lib().orc_glibc_rand(seed, n, _i(out))
    return out


def random_shuffle(a, seed=1):
    """libstdc++ std::random_shuffle of the int32 array a (a copy is returned),
    drawing from glibc rand() after srand(seed), as utility.cpp:679 does on the
    tail of piv."""
    a = np.array(a, np.int32, copy=True)
    lib().orc_random_shuffle(_i(a), a.size, seed)
    return a


def set_scalar(on):
    """Select the scalar (1) or the 4-wide vector (0, default) form of the
    K-means distance loops; both give the same labels and piv."""
    lib().orc_set_scalar(int(bool(on)))


def gather(psi, piv, rk):
    psi = _psi_arr(psi)
    nstate, ng = psi.shape
    piv = np.ascontiguousarray(piv, np.int32)
    out = np.empty((rk, nstate), np.float64)  # column-major nstate x rk
    lib().orc_gather(_d(psi), ng, nstate, _i(piv), rk, _d(out))
    return out


def zct_rows(psi, psimu, nocc, nv1, nc1, nv2, nc2, r0=0, r1=None):
    """Rows [r0, r1) of Z C^T; returns (rk, r1-r0) array (column-major rows x rk)."""
    psi = _psi_arr(psi)
    nstate, ng = psi.shape
    r1 = ng if r1 is None else r1
    psimu = np.ascontiguousarray(psimu, np.float64)
    rk = psimu.shape[0]
    out = np.empty((rk, r1 - r0), np.float64)
    lib().orc_zct_rows(_d(psi), ng, nstate, _d(psimu), rk, nocc, nv1, nc1, nv2, nc2,
                       r0, r1, _d(out), r1 - r0)
    return out


def zct_pick(psi, psimu, nocc, nv1, nc1, nv2, nc2, rows):
    """Z C^T at the given grid rows; returns (rk, len(rows)) (column-major rows x rk)."""
    psi = _psi_arr(psi)
    nstate, ng = psi.shape
    psimu = np.ascontiguousarray(psimu, np.float64)
    rk = psimu.shape[0]
    rows = np.ascontiguousarray(rows, np.int32)
    out = np.empty((rk, rows.size), np.float64)
    lib().orc_zct_pick(_d(psi), ng, nstate, _d(psimu), rk, nocc, nv1, nc1, nv2, nc2, _i(rows),
                       rows.size, _d(out), rows.size)
    return out


def cct_from_psimu(psi, psimu, piv, nocc, nv1, nc1, nv2, nc2):
    """C C^T + eps I from the oracle's Z C^T rows at the pivots
    (src/cpu/getbasis_scalapack.cpp:390-428).  Returns (P (rk, rk) column-major, eps)."""
    rk = psimu.shape[0]
    rows = np.asarray(piv[:rk], np.int32)
    xi_rows = zct_pick(psi, psimu, nocc, nv1, nc1, nv2, nc2, rows)  # (rk cols, rk rows)
    P = np.ascontiguousarray(xi_rows)  # P[nu][mu] = Xi(piv[mu], nu): column-major P(mu, nu)
    tr = 0.0
    for nu in range(rk):
        tr += P[nu, nu]
    eps = 1e-10 * tr / float(rk)
    for nu in range(rk):
        P[nu, nu] += eps
    return P, eps


def cct(xi, piv, rk):
    """xi: (rk, Ng) array (column-major Ng x rk).  Returns (P (rk,rk) col-major, eps)."""
    xi = np.ascontiguousarray(xi, np.float64)
    P = np.empty((rk, rk), np.float64)
    eps = lib().orc_cct(_d(xi), xi.shape[1], _i(np.ascontiguousarray(piv, np.int32)), rk, _d(P))
    return P, eps


def potrf(P):
    """Lower Cholesky of a column-major rk x rk array (given as (rk, rk) C array
    holding columns).  Returns (L_colmajor, info)."""
    A = np.array(P, np.float64, copy=True)
    info = lib().orc_potrf(_d(A), A.shape[0])
    return A, info


def potrs_rows(Lcm, B):
    """B: (rk, rows) array (column-major rows x rk), solved in place: B <- B (L L^T)^-1."""
    B = np.array(B, np.float64, copy=True)
    rk, rows = B.shape
    lib().orc_potrs_rows(_d(np.ascontiguousarray(Lcm)), rk, _d(B), rows, rows)
    return B


def fit_error(psi, nocc, nv1, nc1, nv2, nc2, theta, piv):
    psi = _psi_arr(psi)
    nstate, ng = psi.shape
    theta = np.ascontiguousarray(theta, np.float64)
    rk = theta.shape[0]
    znorm = np.zeros(1, np.float64)
    e = lib().orc_fit_error(_d(psi), ng, ng, nstate, nocc, nv1, nc1, nv2, nc2, _d(theta), ng, rk,
                            _i(np.ascontiguousarray(piv, np.int32)), _d(znorm))
    return e


def isdf(psi, length, ngrid, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, max_iter):
    """Whole CPU ISDF (scalapack + Kmeans semantics).  Returns dict(theta (rk, Ng),
    piv, rk, rk_kmeans, iters, trace, status)."""
    psi = _psi_arr(psi)
    ng = psi.shape[1]
    rk = irank(mu_points, nv1, nc1, nv2, nc2)
    theta = np.zeros((rk, ng), np.float64)
    piv = np.zeros(ng, np.int32)
    rkk = np.zeros(1, np.int32)
    iters = np.zeros(1, np.int32)
    trace = np.zeros(max(max_iter, 1), np.int32)
    L = np.asarray(length, np.float64)
    G = np.asarray(ngrid, np.int32)
    st = lib().orc_isdf(_d(psi), ng, _d(L), _i(G), nocc, nstate, nv1, nc1, nv2, nc2, mu_points,
                        _d(theta), _i(piv), max_iter, _i(rkk), _i(iters), _i(trace))
    it = int(iters[0])
    return dict(theta=theta, piv=piv, rk=rk, rk_kmeans=int(rkk[0]), iters=it,
                trace=trace[:it].copy(), status=st)


# ---------------------------------------------------------------------------
# FFT Coulomb check, src/common/checkcode.cpp:168-312 with src/common/fourier.cpp.
# ---------------------------------------------------------------------------

def gkk(length, ngrid):
    """|G|^2/2 on the full FFT grid in grid order (x fastest), fourier.cpp:55-94."""
    ks = []
    for d in range(3):
        n = ngrid[d]
        k = np.empty(n)
        for i in range(n // 2 + 1):
            k[i] = i * 2.0 * np.pi / length[d]
        for i in range(n // 2 + 1, n):
            k[i] = (i - n) * 2.0 * np.pi / length[d]
        ks.append(k)
    kx, ky, kz = ks
    g = (kx[None, None, :] * kx[None, None, :] + ky[None, :, None] * ky[None, :, None]
         + kz[:, None, None] * kz[:, None, None]) / 2.0
    return g.reshape(-1)


def coulomb_kernel(length, ngrid, radius=5.0):
    """Truncated Coulomb kernel Vcoul(G), checkcode.cpp:168-182."""
    g = gkk(length, ngrid)
    v = np.empty_like(g)
    small = g < 1e-8
    v[small] = 4.0 * np.pi * radius * radius
    gg = g[~small]
    v[~small] = 8.0 * np.pi / (2.0 * gg) * (1 - np.cos(np.sqrt(2.0 * gg) * radius))
    return v


def _fft_cols(cols, ngrid, vol):
    """Forward FFT of real columns (rows of ``cols``), FFTW sign convention
    (e^{-i}), scaled by vol/Ng as FFTWExecute does (fourier.cpp:191-196); the
    full spectrum, which FFTR2C rebuilds by Hermitian symmetry (checkcode.cpp:9-53)."""
    n0, n1, n2 = ngrid
    a = cols.reshape(cols.shape[0], n2, n1, n0)
    f = np.fft.fftn(a, axes=(1, 2, 3)) * (vol / float(n0 * n1 * n2))
    return f.reshape(cols.shape[0], -1)


def coulomb_error(psi, length, ngrid, nocc, nv1, nc1, nv2, nc2, theta, piv, radius=5.0):
    """Relative Coulomb error ||ijVkl - C^T M C||_F / ||ijVkl||_F, checkcode.cpp:168-312.
    Memory ~ Ng*K1*K2 complex: small problems only."""
    psi = _psi_arr(psi)
    theta = np.ascontiguousarray(theta, np.float64)
    rk = theta.shape[0]
    s1, K1 = nocc - nv1, nv1 + nc1
    s2, K2 = nocc - nv2, nv2 + nc2
    vol = float(length[0] * length[1] * length[2])
    V = coulomb_kernel(length, ngrid, radius)
    # pair columns, (k-s1)*K2 + (j-s2) ordering (checkcode.cpp:64-73)
    Z = (psi[s1:s1 + K1, None, :] * psi[None, s2:s2 + K2, :]).reshape(K1 * K2, -1)
    Zg = _fft_cols(Z, ngrid, vol)
    ijVkl = ((Zg.conj() * V[None, :]) @ Zg.T).real  # PsiPhi^H V PsiPhi (:213-242)
    piv = np.asarray(piv)[:rk]
    Cm = (psi[s1:s1 + K1, None, :][:, :, piv] * psi[None, s2:s2 + K2, :][:, :, piv]).reshape(K1 * K2, rk)
    Tg = _fft_cols(theta, ngrid, vol)
    M = ((Tg.conj() * V[None, :]) @ Tg.T).real  # (:249-288)
    CMC = Cm @ M @ Cm.T  # (:291-295)
    return float(np.linalg.norm(ijVkl - CMC) / np.linalg.norm(ijVkl))


# ---------------------------------------------------------------------------
# QRCP point selection, src/cpu/getpoints_QRCP.cpp:11-500 at one rank.
# ---------------------------------------------------------------------------

def _iround(a):
    # include/common/utility.hpp:18-23
    if a > 0:
        return int(a) if a - int(a) < 0.5 else int(a) + 1
    return int(a) if int(a) - a < 0.5 else int(a) - 1


def drand48_stream(seed):
    """POSIX drand48 after srand48(seed): X <- (0x5DEECE66D X + 0xB) mod 2^48,
    each value X / 2^48 (src/common/isdf.cpp:107; utility.hpp:455-457)."""
    x = ((seed & 0xFFFFFFFF) << 16) | 0x330E
    while True:
        x = (0x5DEECE66D * x + 0xB) & ((1 << 48) - 1)
        yield x / float(1 << 48)


def gaussian_sketch(K1, K2, numpre, seed=41):
    """G1 (K1 x numPre) then G2 (K2 x numPre), column-major fill, Box-Muller on
    drand48 pairs (utility.hpp:460-471, getpoints_QRCP.cpp:182-189)."""
    import math
    rng = drand48_stream(seed)
    out = []
    for K in (K1, K2):
        g = np.empty(K * numpre)
        for t in range(K * numpre):
            a = next(rng)
            b = next(rng)
            g[t] = math.sqrt(-2.0 * math.log(a)) * math.cos(2.0 * 3.14159265358979323846 * b)
        out.append(g.reshape(numpre, K).T.copy())  # column-major (K x numPre)
    return out


def qrcp_points(psi, nocc, nv1, nc1, nv2, nc2, rk):
    """getpoints_QRCP at one rank: sketch, Orth (the left singular vectors,
    lapack::Orth = dgesvd), MG(i + j numPre, r) = phiG(r, i) psiG(r, j), the
    columns above 1e-5 of the largest squared norm, and the column-pivoted QR
    (here LAPACK dgeqp3 through scipy; the reference calls ScaLAPACK pdgeqpf).
    Returns (piv: Ng ints, the kept columns in pivot order then zeros;
    numPre; number of kept columns).  Small problems only (MG is formed)."""
    import scipy.linalg
    psi = _psi_arr(psi)
    nstate, ng = psi.shape
    s1, K1 = nocc - nv1, nv1 + nc1
    s2, K2 = nocc - nv2, nv2 + nc2
    numpre = _iround(np.sqrt(rk * 1.5))
    G1, G2 = gaussian_sketch(K1, K2, numpre)
    G1 = np.linalg.svd(G1, full_matrices=False)[0]
    G2 = np.linalg.svd(G2, full_matrices=False)[0]
    psiG = psi[s1:s1 + K1].T @ G1      # Ng x numPre  (getpoints_QRCP.cpp:200-202)
    phiG = psi[s2:s2 + K2].T @ G2      # (:204-206)
    MG = (phiG.T[None, :, :] * psiG.T[:, None, :]).reshape(numpre * numpre, ng)  # row i + j np
    nrm = (MG * MG).sum(axis=0)
    idx = np.nonzero(nrm > nrm.max() * 1e-5)[0]
    _, _, perm = scipy.linalg.qr(MG[:, idx], pivoting=True, mode="economic")
    piv = np.zeros(ng, np.int32)
    piv[:len(idx)] = idx[perm]
    return piv, numpre, len(idx)


# --- BLAS-3 forms of the basis stages (CPU baseline timing) -------------------
# The reference runs these stages through MKL's pdgemm, pdpotrf and pdtrsm
# (src/cpu/getbasis_scalapack.cpp:297-369, :455, :469-478).  The C++
# restatement above follows netlib loop orders, which pins results but runs at
# scalar speed; these forms call the blocked, threaded BLAS/LAPACK that numpy
# and scipy link (OpenBLAS), so the CPU baseline reflects what a BLAS-3 CPU
# ISDF costs.  They agree with the restatement to rounding (tests/test_oracle.py).

def zct_rows_blas(psi, psimu, nocc, nv1, nc1, nv2, nc2, rows):
    """Z C^T at the given grid rows as dgemm(s) plus the Hadamard product;
    returns (len(rows), rk) row-major (the transpose of zct_pick's layout)."""
    s1, K1, s2, K2 = nocc - nv1, nv1 + nc1, nocc - nv2, nv2 + nc2
    A1 = psi[s1:s1 + K1][:, rows].T
    P1 = A1 @ psimu[:, s1:s1 + K1].T
    if nv1 == nv2 and nc1 == nc2:
        return P1 * P1
    A2 = psi[s2:s2 + K2][:, rows].T
    return P1 * (A2 @ psimu[:, s2:s2 + K2].T)


def potrf_lapack(P):
    """Lower Cholesky factor of the symmetric P (LAPACK dpotrf)."""
    import scipy.linalg as sl
    return sl.cholesky(P, lower=True, check_finite=False)


def potrs_rows_blas(L, Z):
    """Rows of Z (rows x rk, row-major) times (L L^T)^{-1}: the two right-side
    triangular solves of getbasis_scalapack.cpp:469-478 as dtrsm calls."""
    import scipy.linalg as sl
    X = sl.solve_triangular(L, Z.T, lower=True, check_finite=False)
    return sl.solve_triangular(L, X, lower=True, trans="T", check_finite=False).T


def blas_threads():
    """Threads of the BLAS numpy uses (threadpoolctl), or None."""
    try:
        import threadpoolctl
        info = threadpoolctl.threadpool_info()
        n = [i["num_threads"] for i in info if i.get("user_api") == "blas"]
        return max(n) if n else None
    except Exception:
        return None
