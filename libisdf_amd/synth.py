"""Seeded synthetic wavefunctions for the ISDF workloads.

psi_j(r) = sum_G c_jG cos(G.r + phi_jG) over the reciprocal vectors of a cubic
cell with |G|^2/2 < Ecut, c_jG ~ N(0,1) exp(-|G|^2/damp), phi_jG ~ U[0, 2pi).
The grid is n^3 with x fastest: grid index g = ix + n*(iy + n*iz) at
r = (ix, iy, iz) L/n, the order LibISDF's K-means coordinates assume
(src/common/utility.cpp:711-713) and its FFT plans use (src/common/fourier.cpp:43-54).

psi is evaluated exactly by one inverse FFT per state (c e^{i phi} placed at
G; |h| stays below n/2, so nothing aliases), then orthonormalised in FP64 with
CholeskyQR2 so that sum_r psi_i psi_j dV = delta_ij, dV = L^3/Ng.

Random numbers: numpy's PCG64 ``default_rng(seed)``; per state, first the nG
normals, then the nG uniform phases, G in the enumeration order of ``gvectors``.

The result is a (nstate, Ng) float64 C-contiguous array, i.e. the column-major
Ng x nstate matrix the ISDF entry point takes.
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    n: int            # grid points per dimension
    L: float          # cubic cell length (bohr)
    nstate: int
    nocc: int
    nv1: int
    nc1: int
    nv2: int
    nc2: int
    mu_points: int
    kmeans_iters: int
    ecut: float = 8.0
    damp: float = 32.0
    seed: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def ngrid(self):
        return (self.n, self.n, self.n)

    @property
    def length(self):
        return (self.L, self.L, self.L)

    @property
    def ng(self):
        return self.n ** 3


# The five workloads (cfg1..cfg4 single instances; cfg5 is a batch, see batch5()).
WORKLOADS = {
    "cfg1": Workload("cfg1", 32, 10.26, 32, 16, 16, 16, 16, 16, 256, 50, seed=1001),
    "cfg2": Workload("cfg2", 64, 20.52, 128, 64, 64, 64, 64, 64, 1024, 50, seed=1002),
    "cfg3": Workload("cfg3", 96, 30.78, 256, 128, 128, 128, 128, 128, 2048, 50, seed=1003),
    "cfg4": Workload("cfg4", 96, 30.78, 384, 192, 192, 192, 64, 192, 3072, 50,
                     ecut=12.0, damp=48.0, seed=1004),
}


def batch5(count=24, seed=5000):
    """The uneven batch: grid n ~ U{24,32,40,48,64}, L = 0.32 n, nstate ~ U{16..96}
    even, nocc = nstate/2, every pair count = nocc, mu_points = 8 nocc, 30 K-means
    iterations.  Instance i draws its wavefunctions with seed ``seed + 1 + i``."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice([24, 32, 40, 48, 64]))
        nstate = int(rng.choice(np.arange(16, 97, 2)))
        nocc = nstate // 2
        out.append(Workload("cfg5_%02d" % i, n, 0.32 * n, nstate, nocc, nocc, nocc, nocc, nocc,
                            8 * nocc, 30, seed=seed + 1 + i))
    return out


def gvectors(L, ecut):
    """Integer triples (h, k, l) with |2 pi/L (h,k,l)|^2 / 2 < ecut, in lexicographic
    (h, k, l) order, and their |G|^2."""
    b = 2.0 * np.pi / L
    hmax = int(np.floor(np.sqrt(2.0 * ecut) / b)) + 1
    r = np.arange(-hmax, hmax + 1)
    h, k, l = np.meshgrid(r, r, r, indexing="ij")
    g2 = (h * h + k * k + l * l) * (b * b)
    m = g2 / 2.0 < ecut
    return np.stack([h[m], k[m], l[m]], axis=1), g2[m]


def wavefunctions(w: Workload, orthonormalise=True, chunk=16):
    n = w.n
    hkl, g2 = gvectors(w.L, w.ecut)
    assert np.abs(hkl).max() < n // 2, "grid too coarse for the cutoff"
    rng = np.random.default_rng(w.seed)
    damp = np.exp(-g2 / w.damp)
    psi = np.empty((w.nstate, n ** 3), np.float64)
    iz = hkl[:, 2] % n
    iy = hkl[:, 1] % n
    ix = hkl[:, 0] % n
    for s0 in range(0, w.nstate, chunk):
        s1 = min(s0 + chunk, w.nstate)
        A = np.zeros((s1 - s0, n, n, n), np.complex128)
        for s in range(s0, s1):
            c = rng.standard_normal(len(g2)) * damp
            phi = rng.uniform(0.0, 2.0 * np.pi, len(g2))
            np.add.at(A[s - s0], (iz, iy, ix), c * np.exp(1j * phi))
        # psi(r) = Re sum_G a_G e^{+i G.r}: numpy's ifftn carries 1/Ng.
        vals = np.fft.ifftn(A, axes=(1, 2, 3)).real * float(n ** 3)
        psi[s0:s1] = vals.reshape(s1 - s0, -1)
    if orthonormalise:
        cholesky_qr2(psi, w.L ** 3 / n ** 3)
    return psi


def cholesky_qr2(A, dV):
    """In place: rows of A (states) become orthonormal under sum_r a_i a_j dV."""
    import scipy.linalg as sl
    for _ in range(2):
        S = (A @ A.T) * dV
        Lc = np.linalg.cholesky(S)
        A[...] = sl.solve_triangular(Lc, A, lower=True)
    return A
