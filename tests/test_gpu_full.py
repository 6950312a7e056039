"""Parity at the full sizes of the 96^3 workloads and of the 24-instance batch.

* 96^3, nstate 256, all pair counts 128, rk 2048 (one pair set), FP64:
  weights, the changed-label trace, the final labels and piv bit for bit
  against the oracle's K-means on the same input (50 iterations,
  src/common/utility.cpp:631-988); C C^T + eps I within 1e-13 of the oracle's
  (getbasis_scalapack.cpp:390-428); Theta on 1500 grid rows sampled on the
  device against the oracle's row solves (getbasis_scalapack.cpp:469-478),
  within max(1e-10, 20 kappa u).
* The same input in FP32, BF16X3 and FP32 + one FP64 correction step: piv
  equal to FP64's, Theta within max(2e-6, 0.5 kappa 2^-24) of FP64 (the
  relative Frobenius norm is formed on the device, in row chunks).
* 96^3, nstate 384, pair sets [0,384) x [128,384), rk 3072 (the second GEMM
  of getbasis_scalapack.cpp:319-349), FP64 as above and BF16X3 against FP64.
* The 24 uneven instances (grids 24^3..64^3, nstate 16..96, 30 iterations)
  through one batched call in FP32 and BF16X3: piv of every instance bit for
  bit the oracle's, Theta bit for bit the single solve's of that precision.

Theta at 96^3 is 14.5 GB (21.7 GB for rk 3072): it never leaves the device
whole; only sampled rows and the pivot rows are copied.
"""
import numpy as np
import pytest

import libisdf_amd as isdf
import oracle
from libisdf_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

U = 2.0 ** -53
U32 = 2.0 ** -24


def cond(P):
    ev = np.linalg.eigvalsh(0.5 * (P + P.T))
    return float(ev[-1] / ev[0])


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def dev_rel(a, b, chunk=128):
    """||a - b||_F / ||b||_F for two (rk, Ng) device tensors, in row chunks."""
    num = torch.zeros((), dtype=torch.float64, device=a.device)
    den = torch.zeros((), dtype=torch.float64, device=a.device)
    for r0 in range(0, a.shape[0], chunk):
        d = a[r0:r0 + chunk] - b[r0:r0 + chunk]
        num += (d * d).sum()
        den += (b[r0:r0 + chunk] * b[r0:r0 + chunk]).sum()
    return float(torch.sqrt(num / den))


def solve(w, d_psi, precision, d_theta=None, debug=False):
    rk = isdf.isdf_rank(w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points)
    if d_theta is None:
        d_theta = torch.empty((rk, w.ng), dtype=torch.float64, device="cuda")
    piv = np.full(w.ng, -9, np.int32)
    dbg = None
    if debug:
        dbg = dict(trace=np.zeros(w.kmeans_iters, np.int32),
                   weights=torch.empty(w.ng, dtype=torch.float64, device="cuda"),
                   labels=torch.empty(w.ng, dtype=torch.int32, device="cuda"),
                   cct=torch.empty((rk, rk), dtype=torch.float64, device="cuda"))
    rep = isdf.solve(d_psi, isdf.Domain(w.length, w.ngrid), w.nocc, w.nstate, w.nv1, w.nc1,
                     w.nv2, w.nc2, w.mu_points, d_theta, piv, w.kmeans_iters,
                     precision=precision, debug=dbg)
    torch.cuda.synchronize()
    return d_theta, piv, rep, dbg


class Full:
    """One 96^3 workload: its input, the FP64 solve and the oracle's K-means."""

    def __init__(self, name):
        w = synth.WORKLOADS[name]
        print("[%s] generating psi" % name, flush=True)
        self.w = w
        self.psi = synth.wavefunctions(w)
        self.d_psi = torch.from_numpy(self.psi).cuda()
        print("[%s] FP64 solve" % name, flush=True)
        self.theta, self.piv, self.rep, dbg = solve(w, self.d_psi, isdf.ISDF_PREC_FP64,
                                                    debug=True)
        self.trace = dbg["trace"][:self.rep["kmeans_iters"]].copy()
        self.weights = dbg["weights"].cpu().numpy()
        self.labels = dbg["labels"].cpu().numpy()
        self.cct = dbg["cct"].cpu().numpy()
        print("[%s] oracle K-means (%d iterations)" % (name, w.kmeans_iters), flush=True)
        self.o_weights = oracle.weights(self.psi, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2)
        self.km = oracle.kmeans(self.o_weights, w.mu_points, w.kmeans_iters, w.length, w.ngrid)
        self.rk = w.mu_points

    def release(self):
        del self.theta, self.d_psi
        torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def cfg3():
    f = Full("cfg3")
    yield f
    f.release()


@pytest.fixture(scope="module")
def cfg4():
    f = Full("cfg4")
    yield f
    f.release()


def check_kmeans(f):
    np.testing.assert_array_equal(f.weights.view(np.uint64), f.o_weights.view(np.uint64))
    assert f.rep["npt"] == f.w.ng == f.km["npt"]
    assert f.rep["kmeans_iters"] == f.km["iters"]
    np.testing.assert_array_equal(f.trace, f.km["trace"])
    np.testing.assert_array_equal(f.labels, f.km["labels"])
    np.testing.assert_array_equal(f.piv, f.km["piv"])
    assert f.rep["rk_kmeans"] == f.km["rk_kmeans"]


def check_basis(f, nrows=1500):
    w, rk = f.w, f.rk
    piv = f.km["piv"]
    psimu = oracle.gather(f.psi, piv, rk)
    P, eps = oracle.cct_from_psimu(f.psi, psimu, piv, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2)
    assert f.rep["eps"] == pytest.approx(eps, rel=1e-12)
    assert rel(f.cct, P) < 1e-13
    Lc, info = oracle.potrf(P)
    assert info == 0
    rows = np.sort(np.random.default_rng(11).choice(w.ng, nrows, replace=False)).astype(np.int32)
    B = oracle.zct_pick(f.psi, psimu, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2, rows)
    T = oracle.potrs_rows(Lc, B)
    got = f.theta[:, torch.from_numpy(rows.astype(np.int64)).cuda()].cpu().numpy()
    assert np.isfinite(got).all()
    tol = max(1e-10, 20.0 * cond(P) * U)
    err = rel(got, T)
    print("[%s] Theta rows vs oracle: %.3e (tol %.1e)" % (w.name, err, tol))
    assert err < tol
    # Theta(piv[mu], nu) = delta_{mu nu} up to the eps shift.
    head = f.theta[:, torch.from_numpy(piv[:rk].astype(np.int64)).cuda()].cpu().numpy()
    assert np.abs(head - np.eye(rk)).max() < 1e-6
    return P


def check_mixed(f, precision, P, name):
    d_mix, piv, rep, _ = solve(f.w, f.d_psi, precision)
    np.testing.assert_array_equal(piv, f.piv)
    err = dev_rel(d_mix, f.theta)
    del d_mix
    torch.cuda.empty_cache()
    tol = max(2e-6, 0.5 * cond(P) * U32)
    print("[%s] %s Theta vs FP64: %.3e (tol %.1e)" % (f.w.name, name, err, tol))
    assert err < tol
    return err


def test_cfg3_fp64_kmeans_bitwise(cfg3):
    check_kmeans(cfg3)


def test_cfg3_fp64_cct_and_theta_rows(cfg3):
    cfg3.P = check_basis(cfg3)


@pytest.mark.parametrize("name,prec", [
    ("fp32", isdf.ISDF_PREC_FP32), ("bf16x3", isdf.ISDF_PREC_BF16X3),
    ("fp32+refine", isdf.ISDF_PREC_FP32 | isdf.ISDF_PREC_REFINE)])
def test_cfg3_mixed_against_fp64(cfg3, name, prec):
    P = getattr(cfg3, "P", None)
    if P is None:
        P = cfg3.cct
    check_mixed(cfg3, prec, P, name)


def test_cfg4_fp64_kmeans_bitwise(cfg4):
    assert cfg4.w.nv1 + cfg4.w.nc1 == 384 and cfg4.w.nv2 + cfg4.w.nc2 == 256
    check_kmeans(cfg4)


def test_cfg4_fp64_cct_and_theta_rows(cfg4):
    cfg4.P = check_basis(cfg4)


def test_cfg4_bf16x3_against_fp64(cfg4):
    P = getattr(cfg4, "P", None)
    if P is None:
        P = cfg4.cct
    check_mixed(cfg4, isdf.ISDF_PREC_BF16X3, P, "bf16x3")


@pytest.fixture(scope="module")
def batch24():
    ws = synth.batch5()
    inst, psis = [], []
    print("[batch] generating 24 instances", flush=True)
    for w in ws:
        psi = synth.wavefunctions(w)
        psis.append(psi)
        rk = isdf.isdf_rank(w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points)
        inst.append(dict(d_psi=torch.from_numpy(psi).cuda(), domain=isdf.Domain(w.length, w.ngrid),
                         nocc=w.nocc, nstate=w.nstate, nv1=w.nv1, nc1=w.nc1, nv2=w.nv2,
                         nc2=w.nc2, mu_points=w.mu_points, max_iter=w.kmeans_iters,
                         d_theta=torch.full((rk, w.ng), float("nan"), dtype=torch.float64,
                                            device="cuda"),
                         piv=np.full(w.ng, -4, np.int32)))
    print("[batch] oracle K-means", flush=True)
    kms = []
    for w, psi in zip(ws, psis):
        wt = oracle.weights(psi, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2)
        kms.append(oracle.kmeans(wt, w.mu_points, w.kmeans_iters, w.length, w.ngrid))
    yield ws, inst, kms
    del inst
    torch.cuda.empty_cache()


@pytest.mark.parametrize("prec", [isdf.ISDF_PREC_FP32, isdf.ISDF_PREC_BF16X3])
def test_batch24_pivots_oracle_and_theta_single(batch24, prec):
    ws, inst, kms = batch24
    assert len(inst) == 24
    reps, brep = isdf.solve_batched(inst, precision=prec)
    torch.cuda.synchronize()
    for i, (w, a, km) in enumerate(zip(ws, inst, kms)):
        np.testing.assert_array_equal(a["piv"], km["piv"], err_msg="instance %d" % i)
        assert reps[i]["kmeans_iters"] == km["iters"]
        d_single, piv, rep, _ = solve(w, a["d_psi"], prec)
        np.testing.assert_array_equal(piv, km["piv"])
        got = a["d_theta"]
        assert bool(torch.isfinite(got).all())
        err = dev_rel(got, d_single)
        assert err < 1e-12, (i, err)
        del d_single
    # A replay of the captured graph reproduces the first call.
    first = [a["d_theta"].clone() for a in inst]
    for a in inst:
        a["d_theta"].fill_(0.0)
    reps2, brep2 = isdf.solve_batched(inst, precision=prec)
    torch.cuda.synchronize()
    assert brep2["graph_replayed"] == 1
    for a, f in zip(inst, first):
        assert bool(torch.equal(a["d_theta"], f))
