"""Parity of the HIP ISDF path (through the C ABI) with the CPU oracle.

K-means outputs (weights, labels, trace, piv) must match bit for bit.  Theta is
compared in relative Frobenius norm with the tolerance stated in each test;
the oracle's C C^T Cholesky is computed from its own Z C^T, so the Theta check
covers the whole FP64 basis chain.
"""
import numpy as np
import pytest

import libisdf_amd as isdf
import oracle
from libisdf_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def gpu_solve(psi, length, ngrid, nocc, nstate, nv1, nc1, nv2, nc2, mu, max_iter):
    ng = psi.shape[1]
    rk = isdf.isdf_rank(nv1, nc1, nv2, nc2, mu)
    d_psi = torch.from_numpy(np.ascontiguousarray(psi)).cuda()
    d_theta = torch.full((rk, ng), float("nan"), dtype=torch.float64, device="cuda")
    piv = np.full(ng, -7, np.int32)
    trace = np.zeros(max(max_iter, 1), np.int32)
    dbg = dict(trace=trace,
               weights=torch.empty(ng, dtype=torch.float64, device="cuda"),
               labels=torch.empty(ng, dtype=torch.int32, device="cuda"),
               cct=torch.empty((rk, rk), dtype=torch.float64, device="cuda"))
    dom = isdf.Domain(length, ngrid)
    rep = isdf.solve(d_psi, dom, nocc, nstate, nv1, nc1, nv2, nc2, mu, d_theta, piv, max_iter,
                     debug=dbg)
    torch.cuda.synchronize()
    return dict(theta=d_theta.cpu().numpy(), piv=piv, rep=rep,
                trace=trace[:rep["kmeans_iters"]].copy(), weights=dbg["weights"].cpu().numpy(),
                labels=dbg["labels"].cpu().numpy(), cct=dbg["cct"].cpu().numpy())


def oracle_solve(psi, length, ngrid, nocc, nstate, nv1, nc1, nv2, nc2, mu, max_iter):
    rk = oracle.irank(mu, nv1, nc1, nv2, nc2)
    w = oracle.weights(psi, nocc, nv1, nc1, nv2, nc2)
    km = oracle.kmeans(w, rk, max_iter, length, ngrid)
    r = oracle.isdf(psi, length, ngrid, nocc, nstate, nv1, nc1, nv2, nc2, mu, max_iter)
    psimu = oracle.gather(psi, r["piv"], rk)
    P, eps = oracle.cct_from_psimu(psi, psimu, r["piv"], nocc, nv1, nc1, nv2, nc2)
    return dict(theta=r["theta"], piv=r["piv"], trace=r["trace"], weights=w, km=km, cct=P,
                eps=eps, rk=rk, rk_kmeans=r["rk_kmeans"])


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


U = 2.0 ** -53


def cond(P):
    ev = np.linalg.eigvalsh(0.5 * (P + P.T))
    return float(ev[-1] / ev[0])


def compare(g, o, theta_tol=None, cct_tol=1e-13):
    """K-means outputs bit for bit; C C^T to FP64 rounding; Theta to theta_tol.
    Two correct FP64 solvers of Theta (C C^T + eps I) = Z C^T agree to about
    kappa(C C^T) u in relative norm, so the default tolerance is 20 kappa u, and
    never looser than that nor tighter than 1e-10 would need."""
    np.testing.assert_array_equal(g["weights"].view(np.uint64), o["weights"].view(np.uint64))
    np.testing.assert_array_equal(g["trace"], o["trace"])
    np.testing.assert_array_equal(g["labels"], o["km"]["labels"])
    np.testing.assert_array_equal(g["piv"], o["piv"])
    assert g["rep"]["rk_kmeans"] == o["rk_kmeans"]
    assert g["rep"]["eps"] == pytest.approx(o["eps"], rel=1e-12)
    # C C^T + eps I: column-major (rk, rk) in both
    assert rel(g["cct"], o["cct"]) < cct_tol
    assert np.isfinite(g["theta"]).all()
    if theta_tol is None:
        theta_tol = max(1e-10, 20.0 * cond(o["cct"]) * U)
    assert rel(g["theta"], o["theta"]) < theta_tol
    return theta_tol


def test_si8_matches_reference_and_oracle(si8):
    psi, ref = si8
    a = ref["args"]
    args = ((10.216,) * 3, (16, 16, 16), a["nocc"], a["nstate"], a["nv1"], a["nc1"], a["nv2"],
            a["nc2"], a["mu_points"], a["kmeans_max_iter"])
    g = gpu_solve(psi, *args)
    assert g["trace"].tolist() == ref["kmeans_trace"]
    o = oracle_solve(psi, *args)
    # C C^T is rank-deficient here (16 states: pair rank 136 < rk = 192, kappa
    # ~ 2e11 after the eps shift), so Theta agrees only to ~kappa u; the fit
    # and Coulomb errors are the well-posed outputs and match the reference.
    tol = compare(g, o)
    assert tol < 1e-3
    dom = isdf.Domain((10.216,) * 3, (16, 16, 16))
    c = isdf.check(psi, g["theta"], g["piv"], dom, a["nocc"], a["nstate"], a["nv1"], a["nc1"],
                   a["nv2"], a["nc2"], a["mu_points"])
    assert float("%.6g" % c["fit_error"]) == ref["fit_error"]
    assert c["coulomb_error"] == pytest.approx(ref["coulomb_error"], rel=1e-4)


def test_cfg1_full_parity():
    w = synth.WORKLOADS["cfg1"]
    psi = synth.wavefunctions(w)
    args = (w.length, w.ngrid, w.nocc, w.nstate, w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points,
            w.kmeans_iters)
    g = gpu_solve(psi, *args)
    o = oracle_solve(psi, *args)
    assert compare(g, o) == 1e-10
    fg = oracle.fit_error(psi, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2, g["theta"], g["piv"])
    fo = oracle.fit_error(psi, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2, o["theta"], o["piv"])
    assert fg == pytest.approx(fo, rel=1e-10)


@pytest.mark.parametrize("case", [
    # (grid, L, nstate, nocc, nv1, nc1, nv2, nc2, mu, iters)  ragged tiles, dual pair sets
    ((10, 12, 9), (4.1, 5.0, 3.7), 22, 11, 9, 7, 5, 11, 100, 40),
    ((16, 14, 18), (6.0, 5.5, 7.0), 40, 20, 20, 20, 8, 20, 300, 30),
    ((12, 12, 12), (5.0, 5.0, 5.0), 24, 12, 12, 12, 12, 12, 0, 25),
])
def test_small_synthetic_parity(case):
    grid, L, nstate, nocc, nv1, nc1, nv2, nc2, mu, iters = case
    rng = np.random.default_rng(77)
    ng = grid[0] * grid[1] * grid[2]
    psi = rng.standard_normal((nstate, ng))
    psi = synth.cholesky_qr2(psi, L[0] * L[1] * L[2] / ng)
    args = (L, grid, nocc, nstate, nv1, nc1, nv2, nc2, mu, iters)
    g = gpu_solve(psi, *args)
    o = oracle_solve(psi, *args)
    compare(g, o, cct_tol=1e-12)


def test_partial_density_sorted_path_and_empty_clusters():
    # psi vanishes on the two lowest z planes (grid point 0 included), so npt <
    # Ng: the K-means points are the threshold list, the tail reads grid point
    # 0 with zero weight (utility.cpp:708), zero-mass clusters are dropped
    # (:834-845), and the centroid sums take the sorted path.
    grid, L = (12, 10, 14), (5.0, 4.0, 6.0)
    ng = grid[0] * grid[1] * grid[2]
    rng = np.random.default_rng(21)
    psi = rng.standard_normal((16, ng))
    psi = synth.cholesky_qr2(psi, L[0] * L[1] * L[2] / ng)
    psi[:, : 2 * grid[0] * grid[1]] = 0.0
    args = (L, grid, 8, 16, 8, 8, 8, 8, 120, 40)
    g = gpu_solve(psi, *args)
    o = oracle_solve(psi, *args)
    assert o["km"]["npt"] == ng - 2 * grid[0] * grid[1]
    assert g["rep"]["npt"] == o["km"]["npt"]
    compare(g, o)


def test_kmeans_maxiter_zero():
    w = synth.WORKLOADS["cfg1"]
    psi = synth.wavefunctions(w)
    args = (w.length, w.ngrid, w.nocc, w.nstate, w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points, 0)
    g = gpu_solve(psi, *args)
    o = oracle_solve(psi, *args)
    assert g["rep"]["kmeans_iters"] == 0
    # The initial centroids are the first rk grid points, so the pivots sit in
    # one corner and kappa(C C^T) ~ 2e11: Theta agrees to ~kappa u, the fit
    # error to 1e-8.
    compare(g, o)
    fg = oracle.fit_error(psi, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2, g["theta"], g["piv"])
    fo = oracle.fit_error(psi, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2, o["theta"], o["piv"])
    assert fg == pytest.approx(fo, rel=1e-8)


def test_kmeans_rk_above_register_update():
    # rk 4500 on a 24^3 grid: more clusters than the control + update launch
    # keeps in registers (4 per thread of 1024, rk <= 4096), so every update
    # takes the general path.  The K-means outputs must still be the oracle's
    # bit for bit (the full Theta oracle at this rk is too slow on the host).
    grid, L = (24, 24, 24), (7.68, 7.68, 7.68)
    ng = grid[0] * grid[1] * grid[2]
    nstate, nocc, rk, iters = 96, 48, 4500, 10
    psi = synth.cholesky_qr2(np.random.default_rng(4500).standard_normal((nstate, ng)),
                             L[0] * L[1] * L[2] / ng)
    g = gpu_solve(psi, L, grid, nocc, nstate, nocc, nocc, nocc, nocc, rk, iters)
    w = oracle.weights(psi, nocc, nocc, nocc, nocc, nocc)
    km = oracle.kmeans(w, rk, iters, L, grid)
    np.testing.assert_array_equal(g["weights"].view(np.uint64), w.view(np.uint64))
    np.testing.assert_array_equal(g["trace"], km["trace"])
    np.testing.assert_array_equal(g["labels"], km["labels"])
    np.testing.assert_array_equal(g["piv"], km["piv"])
    assert g["rep"]["rk_kmeans"] == km["rk_kmeans"]
    assert np.isfinite(g["theta"]).all()


def test_sparse_density_degenerate_branch_and_not_spd():
    # psi supported on 20 grid points only: npt = 20 < rk, the reference's
    # npt<rk branch (utility.cpp:667-681) fills piv from the threshold list;
    # C C^T then has zero rows, eps = 0, and the Cholesky must fail with the
    # reference's message.
    ng = 8 * 8 * 8
    rng = np.random.default_rng(5)
    psi = np.zeros((6, ng))
    # support away from the first 20 grid points, so that every pivot the
    # branch produces has psi = 0 and trace(C C^T) = 0
    pts = rng.choice(np.arange(20, ng), 20, replace=False)
    psi[:, pts] = rng.standard_normal((6, 20))
    d_psi = torch.from_numpy(psi).cuda()
    rk = 32
    d_theta = torch.zeros((rk, ng), dtype=torch.float64, device="cuda")
    piv = np.zeros(ng, np.int32)
    dom = isdf.Domain((4.0, 4.0, 4.0), (8, 8, 8))
    with pytest.raises(RuntimeError, match="Matrix is not positive definite."):
        isdf.ISDF(d_psi, dom, 3, 6, 3, 3, 3, 3, rk, d_theta, piv, 10)
    w = oracle.weights(psi, 3, 3, 3, 3, 3)
    km = oracle.kmeans(w, rk, 10, (4.0, 4.0, 4.0), (8, 8, 8))
    assert km["branch"] == 1
    np.testing.assert_array_equal(piv, km["piv"])


def test_repeatable_bitwise():
    w = synth.WORKLOADS["cfg1"]
    psi = synth.wavefunctions(w)
    args = (w.length, w.ngrid, w.nocc, w.nstate, w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points, 10)
    a = gpu_solve(psi, *args)
    b = gpu_solve(psi, *args)
    np.testing.assert_array_equal(a["theta"].view(np.uint64), b["theta"].view(np.uint64))
    np.testing.assert_array_equal(a["piv"], b["piv"])


def test_interpolation_property_cfg2():
    # At full size: K-means bit-exact against the oracle, Theta rows sampled
    # against the oracle's row solves, and Theta(piv[mu], :) ~ e_mu.
    w = synth.WORKLOADS["cfg2"]
    psi = synth.wavefunctions(w)
    args = (w.length, w.ngrid, w.nocc, w.nstate, w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points,
            w.kmeans_iters)
    g = gpu_solve(psi, *args)
    wts = oracle.weights(psi, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2)
    np.testing.assert_array_equal(g["weights"].view(np.uint64), wts.view(np.uint64))
    km = oracle.kmeans(wts, w.mu_points, w.kmeans_iters, w.length, w.ngrid)
    np.testing.assert_array_equal(g["trace"], km["trace"])
    np.testing.assert_array_equal(g["piv"], km["piv"])
    rk = w.mu_points
    psimu = oracle.gather(psi, km["piv"], rk)
    P, eps = oracle.cct_from_psimu(psi, psimu, km["piv"], w.nocc, w.nv1, w.nc1, w.nv2, w.nc2)
    assert rel(g["cct"], P) < 1e-13
    Lc, info = oracle.potrf(P)
    assert info == 0
    rows = np.random.default_rng(0).choice(w.ng, 1500, replace=False).astype(np.int32)
    B = oracle.zct_pick(psi, psimu, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2, rows)
    T = oracle.potrs_rows(Lc, B)
    assert rel(g["theta"][:, rows], T) < 1e-10
    head = g["theta"][:, km["piv"][:rk]]  # Theta(piv[mu], nu) as (nu, mu)
    assert np.abs(head - np.eye(rk)).max() < 1e-6


@pytest.mark.parametrize("tiles", [("2", "10", "11", "13"), ("6", "9", "12"),
                                   ("2/2", "2/5", "2/6", "10/5")])
def test_solve_tiles_bitwise_equal(monkeypatch, tiles):
    # Every triangular-solve tile and every FP64 Z C^T tile (MFMA 16x16x4 f64
    # tiles and the vector-FMA tiles with DPP-broadcast operands) accumulates
    # each element as the same k-ordered chain of fused multiply-adds, so they
    # agree bit for bit; the first of each group is an MFMA tile.  An entry is
    # "solve tile" or "solve tile/Z C^T tile".
    w = synth.WORKLOADS["cfg1"]
    psi = synth.wavefunctions(w)
    args = (w.length, w.ngrid, w.nocc, w.nstate, w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points, 10)
    ref = None
    for entry in tiles:
        parts = entry.split("/")
        monkeypatch.setenv("ISDF_POTRS_TILE", parts[0])
        if len(parts) > 1:
            monkeypatch.setenv("ISDF_ZCT_TILE", parts[1])
        else:
            monkeypatch.delenv("ISDF_ZCT_TILE", raising=False)
        g = gpu_solve(psi, *args)
        if ref is None:
            ref = g
            continue
        np.testing.assert_array_equal(g["piv"], ref["piv"])
        np.testing.assert_array_equal(g["cct"].view(np.uint64), ref["cct"].view(np.uint64),
                                      err_msg="tiles %s" % entry)
        np.testing.assert_array_equal(g["theta"].view(np.uint64), ref["theta"].view(np.uint64),
                                      err_msg="tiles %s" % entry)


def test_dual_zct_tiles_bitwise_equal(monkeypatch):
    # The two-product Z C^T (differing pair sets) on the MFMA tile and on both
    # vector-FMA tiles.
    nstate, nocc, nv1, nc1, nv2, nc2, mu = 24, 12, 12, 12, 8, 8, 256
    grid, L = (32, 32, 32), (10.26,) * 3
    ng = grid[0] * grid[1] * grid[2]
    psi = synth.cholesky_qr2(np.random.default_rng(5).standard_normal((nstate, ng)),
                             L[0] * L[1] * L[2] / ng)
    args = (L, grid, nocc, nstate, nv1, nc1, nv2, nc2, mu, 10)
    ref = None
    for z in ("4", "5", "6"):
        monkeypatch.setenv("ISDF_ZCT_TILE", z)
        g = gpu_solve(psi, *args)
        if ref is None:
            ref = g
            continue
        np.testing.assert_array_equal(g["cct"].view(np.uint64), ref["cct"].view(np.uint64),
                                      err_msg="Z C^T tile %s" % z)
        np.testing.assert_array_equal(g["theta"].view(np.uint64), ref["theta"].view(np.uint64),
                                      err_msg="Z C^T tile %s" % z)


@pytest.mark.parametrize("pad", [33, 128])
def test_strided_psi_solve_bitwise_equal(pad):
    # psi in a wider buffer (ld_psi = Ng + pad, NaN beyond Ng): an odd leading
    # dimension takes the FP64 Z C^T off the vector-FMA tile (its copies need
    # 16-byte aligned columns) onto the MFMA tile, an even one keeps it; both
    # give the bits of the compact buffer.
    w = synth.WORKLOADS["cfg1"]
    psi = synth.wavefunctions(w)
    args = (w.length, w.ngrid, w.nocc, w.nstate, w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points, 10)
    ref = gpu_solve(psi, *args)
    ng, rk = w.ng, w.mu_points
    wide = np.full((w.nstate, ng + pad), np.nan)
    wide[:, :ng] = psi
    d_wide = torch.from_numpy(wide).cuda()
    d_theta = torch.full((rk, ng), float("nan"), dtype=torch.float64, device="cuda")
    piv = np.zeros(ng, np.int32)
    isdf.solve(d_wide, isdf.Domain(w.length, w.ngrid), w.nocc, w.nstate, w.nv1, w.nc1, w.nv2,
               w.nc2, w.mu_points, d_theta, piv, 10, ld_psi=ng + pad)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(piv, ref["piv"])
    np.testing.assert_array_equal(d_theta.cpu().numpy().view(np.uint64),
                                  ref["theta"].view(np.uint64))
