"""QRCP point selection on the device (isdf_problem.select = ISDF_SELECT_QRCP)
against the oracle's restatement of getpoints_QRCP (LAPACK's pivoted QR on the
explicitly formed sketch matrix MG): the first rk pivots exactly, the kept
columns beyond them as a set, and Theta for those pivots within the
conditioning bound of the K-means tests."""
import numpy as np
import pytest

import libisdf_amd as isdf
import oracle
from libisdf_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

U = 2.0 ** -53

CASES = [
    # (grid, L, nstate, nocc, nv1, nc1, nv2, nc2, rk): ragged grid, dual pair sets
    ((10, 9, 8), (5.0, 4.5, 4.0), 20, 10, 8, 6, 7, 5, 60),
    ((16, 16, 16), (8.0, 8.0, 8.0), 40, 20, 10, 10, 10, 10, 160),
]


def make_psi(grid, L, nstate, seed):
    rng = np.random.default_rng(seed)
    ng = grid[0] * grid[1] * grid[2]
    return synth.cholesky_qr2(rng.standard_normal((nstate, ng)), L[0] * L[1] * L[2] / ng)


def gpu_qrcp(psi, grid, L, nocc, nstate, nv1, nc1, nv2, nc2, rk, precision=isdf.ISDF_PREC_FP64):
    ng = psi.shape[1]
    d_psi = torch.from_numpy(np.ascontiguousarray(psi)).cuda()
    d_theta = torch.empty((rk, ng), dtype=torch.float64, device="cuda")
    piv = np.full(ng, -5, np.int32)
    rep = isdf.solve(d_psi, isdf.Domain(L, grid), nocc, nstate, nv1, nc1, nv2, nc2, rk, d_theta,
                     piv, 10, precision=precision, select=isdf.ISDF_SELECT_QRCP)
    torch.cuda.synchronize()
    return d_theta.cpu().numpy(), piv, rep


@pytest.mark.parametrize("case", CASES)
def test_qrcp_pivots_and_theta_match_oracle(case):
    grid, L, nstate, nocc, nv1, nc1, nv2, nc2, rk = case
    print("qrcp case", case, flush=True)
    psi = make_psi(grid, L, nstate, 91)
    theta, piv, rep = gpu_qrcp(psi, grid, L, nocc, nstate, nv1, nc1, nv2, nc2, rk)
    opiv, numpre, nkept = oracle.qrcp_points(psi, nocc, nv1, nc1, nv2, nc2, rk)
    np.testing.assert_array_equal(piv[:rk], opiv[:rk])
    assert set(piv[rk:nkept].tolist()) == set(opiv[rk:nkept].tolist())
    assert (piv[nkept:] == 0).all()
    assert rep["kmeans_iters"] == 0
    # Theta for these pivots: the oracle's Z C^T, C C^T + eps I, Cholesky and solves.
    psimu = oracle.gather(psi, opiv, rk)
    B = oracle.zct_rows(psi, psimu, nocc, nv1, nc1, nv2, nc2)
    P, eps = oracle.cct_from_psimu(psi, psimu, opiv, nocc, nv1, nc1, nv2, nc2)
    Lc, info = oracle.potrf(P)
    assert info == 0
    ot = oracle.potrs_rows(Lc, B)
    ev = np.linalg.eigvalsh(0.5 * (P + P.T))
    tol = max(1e-10, 20.0 * ev[-1] / ev[0] * U)
    assert np.linalg.norm(theta - ot) / np.linalg.norm(ot) < tol
    # the selection does not depend on the arithmetic of the basis
    _, piv3, _ = gpu_qrcp(psi, grid, L, nocc, nstate, nv1, nc1, nv2, nc2, rk,
                          precision=isdf.ISDF_PREC_BF16X3)
    np.testing.assert_array_equal(piv3, piv)


def test_qrcp_host_interface_and_limits():
    grid, L, nstate, nocc, nv1, nc1, nv2, nc2, rk = CASES[0]
    psi = make_psi(grid, L, nstate, 92)
    ng = psi.shape[1]
    _, piv, _ = gpu_qrcp(psi, grid, L, nocc, nstate, nv1, nc1, nv2, nc2, rk)
    theta = np.empty((rk, ng))
    piv2 = np.zeros(ng, np.int32)
    isdf.ISDF_host(np.ascontiguousarray(psi), isdf.Domain(L, grid), nocc, nstate, nv1, nc1, nv2,
                   nc2, rk, theta, piv2, 10, 32, "scalapack", "QRCP")
    np.testing.assert_array_equal(piv2, piv)
    # numPre = IRound(sqrt(1.5 rk)) larger than a pair set: rejected
    with pytest.raises(ValueError, match="exceeds a pair set"):
        gpu_qrcp(psi, grid, L, nocc, nstate, 3, 2, nv2, nc2, 100)


def test_qrcp_vector_and_mfma_updates_agree(monkeypatch):
    """The R(k, .) update runs on FP64 MFMA when q and the sketch tile fit its
    LDS budget, else on the vector FMA kernel (ISDF_QRCP_VALU forces it): the
    same pivots either way on a non-degenerate case."""
    grid, L, nstate, nocc, nv1, nc1, nv2, nc2, rk = CASES[1]
    psi = make_psi(grid, L, nstate, 93)
    _, piv_m, _ = gpu_qrcp(psi, grid, L, nocc, nstate, nv1, nc1, nv2, nc2, rk)
    monkeypatch.setenv("ISDF_QRCP_VALU", "1")
    _, piv_v, _ = gpu_qrcp(psi, grid, L, nocc, nstate, nv1, nc1, nv2, nc2, rk)
    np.testing.assert_array_equal(piv_m, piv_v)


def test_cpp_host_dropin_with_qrcp(tmp_path):
    """The C++ host-pointer ISDF (include/isdf.hpp:39) with s1 = "QRCP", as the
    reference's example calls it, gives the C-ABI QRCP solve's piv and Theta."""
    import os
    import subprocess
    grid, L, nstate, nocc, nv1, nc1, nv2, nc2, rk = CASES[0]
    psi = make_psi(grid, L, nstate, 94)
    ng = psi.shape[1]
    theta, piv, _ = gpu_qrcp(psi, grid, L, nocc, nstate, nv1, nc1, nv2, nc2, rk)
    np.ascontiguousarray(psi).tofile(tmp_path / "psi.bin")
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin", "_build",
                       "caller_host")
    args = [exe, str(tmp_path / "psi.bin"), *map(str, grid), *map(str, L), str(nocc), str(nstate),
            str(nv1), str(nc1), str(nv2), str(nc2), str(rk), "10", "0",
            str(tmp_path / "theta.bin"), str(tmp_path / "piv.bin"), str(tmp_path / "isdf.0"),
            "scalapack", "QRCP", "32"]
    print("drop-in caller qrcp", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    np.testing.assert_array_equal(np.fromfile(tmp_path / "piv.bin", np.int32), piv)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "theta.bin").reshape(rk, ng), theta)
