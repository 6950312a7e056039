"""check=true on the device (isdf_hip_check_device): the fit error and the FFT
Coulomb error of checkcode.cpp:54-313, against the oracle's numpy restatement
(np.fft spectra, checkcode.cpp:168-312), the host validator (direct DFT on the
CPU) and the reference's printed Si8 values.  The oracle and the host validator
take the same Theta and piv, including a Theta with a planted error, so a wrong
sign, index or scale in the device spectra or Gram products shows."""
import numpy as np
import pytest

import libisdf_amd as isdf
import oracle
from libisdf_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def solve(psi, L, grid, nocc, nstate, nv1, nc1, nv2, nc2, mu, iters):
    ng = psi.shape[1]
    rk = isdf.isdf_rank(nv1, nc1, nv2, nc2, mu)
    d_psi = torch.from_numpy(np.ascontiguousarray(psi)).cuda()
    d_theta = torch.empty((rk, ng), dtype=torch.float64, device="cuda")
    piv = np.zeros(ng, np.int32)
    isdf.solve(d_psi, isdf.Domain(L, grid), nocc, nstate, nv1, nc1, nv2, nc2, mu, d_theta, piv,
               iters)
    torch.cuda.synchronize()
    return d_psi, d_theta, piv


def test_si8_device_check_matches_reference_oracle_and_host(si8):
    psi, ref = si8
    a = ref["args"]
    L, G = (10.216,) * 3, (16, 16, 16)
    args = (a["nocc"], a["nstate"], a["nv1"], a["nc1"], a["nv2"], a["nc2"], a["mu_points"])
    print("device check si8", flush=True)
    d_psi, d_theta, piv = solve(psi, L, G, *args, a["kmeans_max_iter"])
    dom = isdf.Domain(L, G)
    c = isdf.check_device(d_psi, d_theta, piv, dom, *args)
    assert float("%.6g" % c["fit_error"]) == ref["fit_error"]
    # Theta is ill-conditioned here (kappa ~ 2e11): the Coulomb error of this
    # Theta agrees with the reference's printed one to ~1e-4.
    assert c["coulomb_error"] == pytest.approx(ref["coulomb_error"], rel=1e-4)
    theta = d_theta.cpu().numpy()
    h = isdf.check(psi, theta, piv, dom, *args)
    o = oracle.coulomb_error(psi, L, G, a["nocc"], a["nv1"], a["nc1"], a["nv2"], a["nc2"], theta,
                             piv)
    assert c["coulomb_error"] == pytest.approx(h["coulomb_error"], rel=1e-6)
    assert c["coulomb_norm"] == pytest.approx(h["coulomb_norm"], rel=1e-12)
    assert c["coulomb_error"] == pytest.approx(o, rel=1e-6)
    assert isdf.check_device(d_psi, d_theta, piv, dom, *args) == c  # deterministic
    f = isdf.check_device(d_psi, d_theta, piv, dom, *args, coulomb=False)
    assert f["fit_error"] == c["fit_error"] and f["coulomb_error"] == -1.0


@pytest.mark.parametrize("case", [
    # (grid, L, nstate, nocc, nv1, nc1, nv2, nc2, mu, iters): odd and unequal axes,
    # dual pair sets, ragged pair and rk tiles
    ((10, 12, 9), (4.1, 5.0, 3.7), 22, 11, 9, 7, 5, 11, 100, 40),
    ((16, 14, 18), (6.0, 5.5, 7.0), 40, 20, 20, 20, 8, 20, 300, 30),
    # radix 11 and 13 FFT axes and a prime axis (17) on the direct DFT
    ((11, 13, 17), (4.3, 5.1, 6.2), 22, 11, 9, 7, 5, 11, 100, 40),
])
def test_device_coulomb_ragged_and_planted_error(case):
    grid, L, nstate, nocc, nv1, nc1, nv2, nc2, mu, iters = case
    print("device check case", case, flush=True)
    rng = np.random.default_rng(79)
    ng = grid[0] * grid[1] * grid[2]
    psi = synth.cholesky_qr2(rng.standard_normal((nstate, ng)), L[0] * L[1] * L[2] / ng)
    args = (nocc, nstate, nv1, nc1, nv2, nc2, mu)
    d_psi, d_theta, piv = solve(psi, L, grid, *args, iters)
    dom = isdf.Domain(L, grid)
    theta = d_theta.cpu().numpy()
    bad = theta.copy()
    bad[mu // 2, ::5] -= 0.1 * np.abs(theta).max()
    for th in (theta, bad):
        d_th = torch.from_numpy(th).cuda()
        c = isdf.check_device(d_psi, d_th, piv, dom, *args)
        o = oracle.coulomb_error(psi, L, grid, nocc, nv1, nc1, nv2, nc2, th, piv)
        h = isdf.check(psi, th, piv, dom, *args)
        assert c["coulomb_error"] == pytest.approx(o, rel=1e-6)
        assert c["coulomb_error"] == pytest.approx(h["coulomb_error"], rel=1e-6)
        assert c["fit_error"] == pytest.approx(h["fit_error"], rel=1e-12)


def test_device_check_with_strided_psi(si8):
    """ld_psi > Ng (NaN padding beyond Ng): the same numbers as the compact buffer."""
    psi, ref = si8
    a = ref["args"]
    L, G = (10.216,) * 3, (16, 16, 16)
    args = (a["nocc"], a["nstate"], a["nv1"], a["nc1"], a["nv2"], a["nc2"], a["mu_points"])
    d_psi, d_theta, piv = solve(psi, L, G, *args, a["kmeans_max_iter"])
    nstate, ng = psi.shape
    wide = np.full((nstate, ng + 33), np.nan)
    wide[:, :ng] = psi
    dom = isdf.Domain(L, G)
    c = isdf.check_device(d_psi, d_theta, piv, dom, *args)
    cw = isdf.check_device(torch.from_numpy(wide).cuda(), d_theta, piv, dom, *args,
                           ld_psi=ng + 33)
    assert cw == c


def test_python_isdf_check_runs_on_device_cfg1():
    # ISDF(check=True) in Python checks on the device, as the C++ drop-in does,
    # and agrees with the host validator on the same Theta and piv.
    w = synth.WORKLOADS["cfg1"]
    psi = synth.wavefunctions(w)
    rk = w.mu_points
    d_psi = torch.from_numpy(psi).cuda()
    d_theta = torch.empty((rk, w.ng), dtype=torch.float64, device="cuda")
    piv = np.zeros(w.ng, np.int32)
    dom = isdf.Domain(w.length, w.ngrid)
    rep = isdf.ISDF(d_psi, dom, w.nocc, w.nstate, w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points,
                    d_theta, piv, w.kmeans_iters, check=True)
    host = isdf.check(psi, d_theta.cpu().numpy(), piv, dom, w.nocc, w.nstate, w.nv1, w.nc1,
                      w.nv2, w.nc2, w.mu_points)
    assert rep["fit_error"] == pytest.approx(host["fit_error"], rel=1e-10)
    assert rep["coulomb_error"] == pytest.approx(host["coulomb_error"], rel=1e-8)
    assert rep["coulomb_error"] > 0
