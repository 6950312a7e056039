"""The mixed-precision forms (ISDF_PREC_FP32, ISDF_PREC_BF16X3, each optionally
with ISDF_PREC_REFINE) through the C ABI.

They have no counterpart in the reference: K-means is unchanged (piv stays bit
for bit the reference's), and Theta is compared with the FP64 result, which
the FP64 tests pin to the oracle.  Both contractions accumulate in FP32, so
two FP32-accurate solves agree with FP64 to about kappa(C C^T) 2^-24; the
tolerance in each test is max(2e-6, 0.5 kappa 2^-24) for FP32 operands (the
measured ratio err / (kappa 2^-24) is about 0.12 on the 32^3 workload).
"""
import numpy as np
import pytest

import libisdf_amd as isdf
import oracle
from libisdf_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

U32 = 2.0 ** -24
MODES = [isdf.ISDF_PREC_FP32, isdf.ISDF_PREC_BF16X3,
         isdf.ISDF_PREC_FP32 | isdf.ISDF_PREC_REFINE,
         isdf.ISDF_PREC_BF16X3 | isdf.ISDF_PREC_REFINE]


def run(psi, length, ngrid, nocc, nstate, nv1, nc1, nv2, nc2, mu, iters, precision):
    ng = psi.shape[1]
    rk = isdf.isdf_rank(nv1, nc1, nv2, nc2, mu)
    d_psi = torch.from_numpy(np.ascontiguousarray(psi)).cuda()
    d_theta = torch.full((rk, ng), float("nan"), dtype=torch.float64, device="cuda")
    piv = np.full(ng, -3, np.int32)
    cct = torch.empty((rk, rk), dtype=torch.float64, device="cuda")
    rep = isdf.solve(d_psi, isdf.Domain(length, ngrid), nocc, nstate, nv1, nc1, nv2, nc2, mu,
                     d_theta, piv, iters, precision=precision, debug=dict(cct=cct))
    torch.cuda.synchronize()
    return d_theta.cpu().numpy(), piv, rep, cct.cpu().numpy()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cond(P):
    ev = np.linalg.eigvalsh(0.5 * (P + P.T))
    return float(ev[-1] / ev[0])


@pytest.fixture(scope="module")
def cfg1():
    w = synth.WORKLOADS["cfg1"]
    psi = synth.wavefunctions(w)
    args = (w.length, w.ngrid, w.nocc, w.nstate, w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points,
            w.kmeans_iters)
    th64, piv64, rep64, cct64 = run(psi, *args, isdf.ISDF_PREC_FP64)
    fit64 = oracle.fit_error(psi, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2, th64, piv64)
    return w, psi, args, th64, piv64, cct64, fit64


@pytest.mark.parametrize("mode", MODES)
def test_mixed_cfg1_against_fp64(cfg1, mode):
    w, psi, args, th64, piv64, cct64, fit64 = cfg1
    th, piv, rep, cct = run(psi, *args, mode)
    np.testing.assert_array_equal(piv, piv64)
    assert rep["potrf_info"] == 0
    # C C^T from the FP32 contraction: within a few 2^-24 of the FP64 one
    assert rel(cct, cct64) < 8 * U32
    kappa = cond(cct64)
    tol = max(2e-6, 0.5 * kappa * U32)
    err = rel(th, th64)
    assert err < tol, (err, tol, kappa)
    fit = oracle.fit_error(psi, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2, th, piv)
    # fit error at fixed quality: within 1.1x of FP64 (in fact to ~1e-6)
    assert fit == pytest.approx(fit64, rel=1e-6)
    assert fit <= 1.1 * fit64


def test_refine_does_not_lose_accuracy(cfg1):
    w, psi, args, th64, piv64, cct64, fit64 = cfg1
    for base in (isdf.ISDF_PREC_FP32, isdf.ISDF_PREC_BF16X3):
        e0 = rel(run(psi, *args, base)[0], th64)
        e1 = rel(run(psi, *args, base | isdf.ISDF_PREC_REFINE)[0], th64)
        assert e1 <= e0 * 1.05, (base, e0, e1)


@pytest.mark.parametrize("mode", [isdf.ISDF_PREC_FP32, isdf.ISDF_PREC_BF16X3])
@pytest.mark.parametrize("case", [
    # (grid, L, nstate, nocc, nv1, nc1, nv2, nc2, mu, iters): ragged tiles, k and rk not
    # multiples of 32, differing pair sets (two products in the Z C^T epilogue)
    ((10, 12, 9), (4.1, 5.0, 3.7), 22, 11, 9, 7, 5, 11, 100, 40),
    ((16, 14, 18), (6.0, 5.5, 7.0), 40, 20, 20, 20, 8, 20, 300, 30),
    ((20, 20, 20), (8.0, 8.0, 8.0), 48, 24, 20, 24, 12, 16, 0, 30),
])
def test_mixed_ragged_and_dual_against_oracle(case, mode):
    grid, L, nstate, nocc, nv1, nc1, nv2, nc2, mu, iters = case
    rng = np.random.default_rng(77)
    ng = grid[0] * grid[1] * grid[2]
    psi = synth.cholesky_qr2(rng.standard_normal((nstate, ng)), L[0] * L[1] * L[2] / ng)
    args = (L, grid, nocc, nstate, nv1, nc1, nv2, nc2, mu, iters)
    th, piv, rep, cct = run(psi, *args, mode)
    o = oracle.isdf(psi, L, grid, nocc, nstate, nv1, nc1, nv2, nc2, mu, iters)
    np.testing.assert_array_equal(piv, o["piv"])
    rk = oracle.irank(mu, nv1, nc1, nv2, nc2)
    psimu = oracle.gather(psi, o["piv"], rk)
    P, eps = oracle.cct_from_psimu(psi, psimu, o["piv"], nocc, nv1, nc1, nv2, nc2)
    kappa = cond(P)
    tol = max(2e-6, 0.5 * kappa * U32)
    assert rel(th, o["theta"]) < tol, (rel(th, o["theta"]), tol, kappa)


def test_mixed_repeatable_bitwise(cfg1):
    w, psi, args, *_ = cfg1
    for mode in (isdf.ISDF_PREC_FP32, isdf.ISDF_PREC_BF16X3):
        a = run(psi, *args, mode)[0]
        b = run(psi, *args, mode)[0]
        np.testing.assert_array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("refine", [0, isdf.ISDF_PREC_REFINE])
@pytest.mark.parametrize("case", ["cfg1", "ragged"])
def test_bf16x3_theta_tiles_bitwise_equal(cfg1, monkeypatch, case, refine):
    # Every BF16X3 Theta tile runs the same six products per k chunk in the same
    # order on the same bf16 parts (pre-split or split in registers), so the
    # tile choice must not change a bit: 0-2 read pre-split W and FP32 Zt, 3
    # (direct-to-LDS) FP32 W and Zt, 4 FP32 W and a Zt the Z C^T epilogue wrote
    # pre-split (with the correction step it falls back to 3).  The ragged
    # case has partial row and column tiles (rk = 100, Ng = 1080).
    if case == "cfg1":
        psi, args = cfg1[1], cfg1[2]
    else:
        grid, L = (10, 12, 9), (4.1, 5.0, 3.7)
        ng = grid[0] * grid[1] * grid[2]
        psi = synth.cholesky_qr2(np.random.default_rng(77).standard_normal((22, ng)),
                                 L[0] * L[1] * L[2] / ng)
        args = (L, grid, 11, 22, 9, 7, 5, 11, 100, 40)
    out = {}
    for v in (0, 1, 2, 3, 4):
        monkeypatch.setenv("ISDF_LP_THETA_TILE", str(v))
        out[v] = run(psi, *args, isdf.ISDF_PREC_BF16X3 | refine)[0]
    assert np.isfinite(out[0]).all()
    for v in (1, 2, 3, 4):
        np.testing.assert_array_equal(out[v].view(np.uint64), out[0].view(np.uint64))


@pytest.mark.parametrize("case", ["cfg1", "dual"])
def test_bf16x3_zct_tiles_bitwise_equal(cfg1, monkeypatch, case):
    # Z C^T tiles: pre-split operand rows (variants 0-2) or FP32 rows split in
    # registers after direct-to-LDS loads (3: 256 x 128, 4: 256 x 256, which
    # the dual product maps to 3).  The same six products per k chunk in the
    # same order: Theta must not change a bit.  The dual case has two pair
    # sets (two products in the epilogue), rk = 300 and Ng = 4032.
    if case == "cfg1":
        psi, args = cfg1[1], cfg1[2]
    else:
        grid, L = (16, 14, 18), (6.0, 5.5, 7.0)
        ng = grid[0] * grid[1] * grid[2]
        psi = synth.cholesky_qr2(np.random.default_rng(77).standard_normal((40, ng)),
                                 L[0] * L[1] * L[2] / ng)
        args = (L, grid, 20, 40, 20, 20, 8, 20, 300, 30)
    out = {}
    for v in (0, 1, 2, 3, 4):
        monkeypatch.setenv("ISDF_LP_ZCT_TILE", str(v))
        out[v] = run(psi, *args, isdf.ISDF_PREC_BF16X3)[0]
    assert np.isfinite(out[0]).all()
    for v in (1, 2, 3, 4):
        np.testing.assert_array_equal(out[v].view(np.uint64), out[0].view(np.uint64))


def test_mixed_ill_conditioned_raises_not_spd(si8):
    # Si8 has kappa(C C^T) ~ 2e11 (16 states give 136 pair products for rk =
    # 192).  The FP32 contraction's rounding (~1e-7 relative) exceeds the eps
    # shift (1e-10 trace / rk), so C C^T + eps I is not positive definite in
    # the mixed forms: the call must fail with the reference's message, not
    # return a Theta.  The FP64 path handles this case (test_gpu_isdf).
    psi, ref = si8
    a = ref["args"]
    ng = psi.shape[1]
    d_psi = torch.from_numpy(psi).cuda()
    d_theta = torch.empty((a["mu_points"], ng), dtype=torch.float64, device="cuda")
    piv = np.zeros(ng, np.int32)
    dom = isdf.Domain((10.216,) * 3, (16, 16, 16))
    with pytest.raises(RuntimeError, match="Matrix is not positive definite."):
        isdf.solve(d_psi, dom, a["nocc"], a["nstate"], a["nv1"], a["nc1"], a["nv2"], a["nc2"],
                   a["mu_points"], d_theta, piv, a["kmeans_max_iter"],
                   precision=isdf.ISDF_PREC_FP32)
