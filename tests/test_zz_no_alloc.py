"""isdf_hip_solve_noalloc against isdf_hip_solve, through the C ABI.

The no-allocation entry point runs the same kernels on a workspace carved
from a static device array, so Theta and piv must equal the allocating solve's
bit for bit.  Two calls in a row, the second on a smaller problem (the array
keeps the first call's contents), then a mixed-precision call, then a problem
whose workspace exceeds the array (rejected before any device work).
"""
import numpy as np
import pytest

import libisdf_amd as isdf
from libisdf_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def both(psi, L, grid, nocc, nstate, nv1, nc1, nv2, nc2, mu, iters, precision):
    ng = psi.shape[1]
    rk = isdf.isdf_rank(nv1, nc1, nv2, nc2, mu)
    dom = isdf.Domain(L, grid)
    d_psi = torch.from_numpy(np.ascontiguousarray(psi)).cuda()
    out = []
    for fn in (isdf.solve, isdf.solve_noalloc):
        d_theta = torch.full((rk, ng), float("nan"), dtype=torch.float64, device="cuda")
        piv = np.full(ng, -2, np.int32)
        rep = fn(d_psi, dom, nocc, nstate, nv1, nc1, nv2, nc2, mu, d_theta, piv, iters,
                 precision=precision)
        torch.cuda.synchronize()
        out.append((d_theta.cpu().numpy(), piv, rep))
    (t0, p0, r0), (t1, p1, r1) = out
    np.testing.assert_array_equal(p1, p0)
    assert np.isfinite(t1).all()
    np.testing.assert_array_equal(t1.view(np.uint64), t0.view(np.uint64))
    for k in ("rk", "rk_kmeans", "kmeans_iters", "npt", "potrf_info", "eps"):
        assert r1[k] == r0[k], k
    assert r1["ms_total"] == 0.0


def test_noalloc_equals_solve_then_smaller_then_mixed(si8):
    w = synth.WORKLOADS["cfg1"]
    psi = synth.wavefunctions(w)
    print("noalloc: 32^3 FP64", flush=True)
    both(psi, w.length, w.ngrid, w.nocc, w.nstate, w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points,
         w.kmeans_iters, isdf.ISDF_PREC_FP64)
    # Second call, smaller problem: the Si8 example (16^3, rk 192).
    psi8, ref = si8
    a = ref["args"]
    print("noalloc: Si8 FP64", flush=True)
    both(psi8, (10.216,) * 3, (16, 16, 16), a["nocc"], a["nstate"], a["nv1"], a["nc1"],
         a["nv2"], a["nc2"], a["mu_points"], a["kmeans_max_iter"], isdf.ISDF_PREC_FP64)
    print("noalloc: 32^3 BF16X3", flush=True)
    both(psi, w.length, w.ngrid, w.nocc, w.nstate, w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points,
         w.kmeans_iters, isdf.ISDF_PREC_BF16X3)


def test_noalloc_rejects_a_workspace_beyond_the_array():
    # 96^3 BF16X3 needs ~12 GB of workspace; the call fails before any kernel
    # runs, so the buffers are left uninitialised.
    w = synth.WORKLOADS["cfg3"]
    rk = w.mu_points
    d_psi = torch.empty((w.nstate, w.ng), dtype=torch.float64, device="cuda")
    d_theta = torch.empty((rk, w.ng), dtype=torch.float64, device="cuda")
    piv = np.zeros(w.ng, np.int32)
    with pytest.raises(ValueError, match="bytes of device workspace"):
        isdf.solve_noalloc(d_psi, isdf.Domain(w.length, w.ngrid), w.nocc, w.nstate, w.nv1,
                           w.nc1, w.nv2, w.nc2, w.mu_points, d_theta, piv, w.kmeans_iters,
                           precision=isdf.ISDF_PREC_BF16X3)
    del d_psi, d_theta
    torch.cuda.empty_cache()
