"""Device fit error (isdf_hip_fit_error) against the oracle's restatement of
checkcode's fit error (src/common/checkcode.cpp:62-166) and the host
validator, on Theta and piv from the HIP solve.  Cases: the reference's Si8
output (printed digits), ragged tiles with dual pair sets, the pair-slab loop
forced to many slabs, and a Theta with a planted error (the oracle is handed
the same perturbed Theta, so a dropped or misplaced term in the device
residual would show)."""
import numpy as np
import pytest

import libisdf_amd as isdf
import oracle
from libisdf_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def solve_and_fit(psi, L, grid, nocc, nstate, nv1, nc1, nv2, nc2, mu, iters):
    ng = psi.shape[1]
    rk = isdf.isdf_rank(nv1, nc1, nv2, nc2, mu)
    d_psi = torch.from_numpy(np.ascontiguousarray(psi)).cuda()
    d_theta = torch.empty((rk, ng), dtype=torch.float64, device="cuda")
    piv = np.zeros(ng, np.int32)
    dom = isdf.Domain(L, grid)
    isdf.solve(d_psi, dom, nocc, nstate, nv1, nc1, nv2, nc2, mu, d_theta, piv, iters)
    torch.cuda.synchronize()
    fit = isdf.fit_error(d_psi, d_theta, piv, dom, nocc, nstate, nv1, nc1, nv2, nc2, mu)
    return d_psi, d_theta, piv, dom, fit


def test_si8_fit_error_matches_reference_and_oracle(si8):
    psi, ref = si8
    a = ref["args"]
    args = ((10.216,) * 3, (16, 16, 16), a["nocc"], a["nstate"], a["nv1"], a["nc1"], a["nv2"],
            a["nc2"], a["mu_points"], a["kmeans_max_iter"])
    print("fit error si8", flush=True)
    _, d_theta, piv, dom, fit = solve_and_fit(psi, *args)
    assert float("%.6g" % fit["fit_error"]) == ref["fit_error"]
    theta = d_theta.cpu().numpy()
    fo = oracle.fit_error(psi, a["nocc"], a["nv1"], a["nc1"], a["nv2"], a["nc2"], theta, piv)
    assert fit["fit_error"] == pytest.approx(fo, rel=1e-11)
    h = isdf.check(psi, theta, piv, dom, a["nocc"], a["nstate"], a["nv1"], a["nc1"], a["nv2"],
                   a["nc2"], a["mu_points"], coulomb=False)
    assert fit["fit_error"] == pytest.approx(h["fit_error"], rel=1e-12)
    assert fit["z_norm"] == pytest.approx(h["z_norm"], rel=1e-13)


@pytest.mark.parametrize("case", [
    # (grid, L, nstate, nocc, nv1, nc1, nv2, nc2, mu, iters)  ragged rows, pairs and rk
    ((10, 12, 9), (4.1, 5.0, 3.7), 22, 11, 9, 7, 5, 11, 100, 40),
    ((16, 14, 18), (6.0, 5.5, 7.0), 40, 20, 20, 20, 8, 20, 300, 30),
])
def test_ragged_dual_pair_sets_and_slabs(case, monkeypatch):
    grid, L, nstate, nocc, nv1, nc1, nv2, nc2, mu, iters = case
    print("fit error case", case, flush=True)
    rng = np.random.default_rng(78)
    ng = grid[0] * grid[1] * grid[2]
    psi = synth.cholesky_qr2(rng.standard_normal((nstate, ng)), L[0] * L[1] * L[2] / ng)
    d_psi, d_theta, piv, dom, fit = solve_and_fit(psi, L, grid, nocc, nstate, nv1, nc1, nv2,
                                                  nc2, mu, iters)
    theta = d_theta.cpu().numpy()
    fo = oracle.fit_error(psi, nocc, nv1, nc1, nv2, nc2, theta, piv)
    assert fit["fit_error"] == pytest.approx(fo, rel=1e-11)
    # one pair tile per slab: the slab loop and its partial-slot accumulation
    monkeypatch.setenv("ISDF_FIT_SLAB_COLS", "128")
    f2 = isdf.fit_error(d_psi, d_theta, piv, dom, nocc, nstate, nv1, nc1, nv2, nc2, mu)
    assert f2["fit_error"] == pytest.approx(fit["fit_error"], rel=1e-13)
    # planted error in one Theta column (rows strided so it crosses row strips)
    # (Theta is the least-squares fit, so any change raises the error; the
    # oracle on the same perturbed Theta fixes by how much.)
    bad = theta.copy()
    bad[mu // 3, ::7] += 0.05 * np.abs(theta).max()
    d_bad = torch.from_numpy(bad).cuda()
    f3 = isdf.fit_error(d_psi, d_bad, piv, dom, nocc, nstate, nv1, nc1, nv2, nc2, mu)
    fo3 = oracle.fit_error(psi, nocc, nv1, nc1, nv2, nc2, bad, piv)
    assert f3["fit_error"] > fit["fit_error"] * (1 + 1e-6)
    assert f3["fit_error"] == pytest.approx(fo3, rel=1e-11)


def test_fit_error_deterministic_cfg1():
    w = synth.WORKLOADS["cfg1"]
    psi = synth.wavefunctions(w)
    print("fit error cfg1", flush=True)
    d_psi, d_theta, piv, dom, fit = solve_and_fit(
        psi, w.length, w.ngrid, w.nocc, w.nstate, w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points,
        w.kmeans_iters)
    again = isdf.fit_error(d_psi, d_theta, piv, dom, w.nocc, w.nstate, w.nv1, w.nc1, w.nv2,
                           w.nc2, w.mu_points)
    assert again == fit
    theta = d_theta.cpu().numpy()
    fo = oracle.fit_error(psi, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2, theta, piv)
    assert fit["fit_error"] == pytest.approx(fo, rel=1e-11)


def test_fit_error_and_host_solve_with_strided_psi(si8):
    """ld_psi > Ng: psi rows sit in a wider buffer (isdf_problem.ld_psi), for the
    device fit error and for the host-buffer solve."""
    psi, ref = si8
    a = ref["args"]
    nstate, ng = psi.shape
    L, G, rk = (10.216,) * 3, (16, 16, 16), a["mu_points"]
    dom = isdf.Domain(L, G)
    wide = np.full((nstate, ng + 77), np.nan)
    wide[:, :ng] = psi
    d_psi = torch.from_numpy(np.ascontiguousarray(psi)).cuda()
    d_wide = torch.from_numpy(wide).cuda()
    d_theta = torch.empty((rk, ng), dtype=torch.float64, device="cuda")
    piv = np.zeros(ng, np.int32)
    isdf.solve(d_psi, dom, a["nocc"], nstate, a["nv1"], a["nc1"], a["nv2"], a["nc2"], rk, d_theta,
               piv, a["kmeans_max_iter"])
    torch.cuda.synchronize()
    f = isdf.fit_error(d_psi, d_theta, piv, dom, a["nocc"], nstate, a["nv1"], a["nc1"], a["nv2"],
                       a["nc2"], rk)
    fw = isdf.fit_error(d_wide, d_theta, piv, dom, a["nocc"], nstate, a["nv1"], a["nc1"],
                        a["nv2"], a["nc2"], rk, ld_psi=ng + 77)
    assert fw == f
    L_ = isdf._lib.lib()
    p = isdf._problem(dom, a["nocc"], nstate, a["nv1"], a["nc1"], a["nv2"], a["nc2"], rk,
                      a["kmeans_max_iter"], isdf.ISDF_PREC_FP64, ng + 77)
    th = np.empty((rk, ng))
    piv2 = np.zeros(ng, np.int32)
    import ctypes
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L_.isdf_hip_solve_host(ctypes.byref(p), wide.ctypes.data_as(dp), th.ctypes.data_as(dp),
                                piv2.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None, None,
                                None)
    assert rc == 0, isdf._lib.last_error()
    assert np.array_equal(piv2, piv) and np.array_equal(th, d_theta.cpu().numpy())
