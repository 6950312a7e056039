"""The batched solve (isdf_hip_solve_batched) through the C ABI.

Each instance of a batch must come out exactly as the single solve of the same
precision gives it: the same K-means (piv bit for bit), the same kernels for
the Cholesky chain, and grouped contractions whose per-element sums run in the
same k order as the single launches.  The captured hipGraph is replayed
several times, with eager work and synchronisations between the replays, and
every replay must reproduce the first call bit for bit.
"""
import numpy as np
import pytest

import libisdf_amd as isdf
import oracle
from libisdf_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def make_batch(seed, count, partial=None):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice([12, 14, 16, 20]))
        grid = (n, n + 2 * (i % 2), n)
        L = tuple(0.32 * g for g in grid)
        nstate = int(rng.choice([8, 12, 16, 20]))
        nocc = nstate // 2
        ng = grid[0] * grid[1] * grid[2]
        psi = synth.cholesky_qr2(rng.standard_normal((nstate, ng)), L[0] * L[1] * L[2] / ng)
        if partial is not None and i == partial:
            psi[:, : 2 * grid[0] * grid[1]] = 0.0  # npt < Ng: the single-solve fallback
        out.append(dict(psi=psi, grid=grid, L=L, nstate=nstate, nocc=nocc, mu=8 * nocc,
                        iters=20))
    return out


def to_instances(batch):
    inst = []
    for b in batch:
        ng = b["psi"].shape[1]
        inst.append(dict(d_psi=torch.from_numpy(b["psi"]).cuda(),
                         domain=isdf.Domain(b["L"], b["grid"]), nocc=b["nocc"],
                         nstate=b["nstate"], nv1=b["nocc"], nc1=b["nocc"], nv2=b["nocc"],
                         nc2=b["nocc"], mu_points=b["mu"],
                         d_theta=torch.full((b["mu"], ng), float("nan"), dtype=torch.float64,
                                            device="cuda"),
                         piv=np.full(ng, -5, np.int32), max_iter=b["iters"]))
    return inst


def single(a, precision):
    d_theta = torch.empty_like(a["d_theta"])
    piv = np.zeros_like(a["piv"])
    rep = isdf.solve(a["d_psi"], a["domain"], a["nocc"], a["nstate"], a["nv1"], a["nc1"],
                     a["nv2"], a["nc2"], a["mu_points"], d_theta, piv, a["max_iter"],
                     precision=precision)
    torch.cuda.synchronize()
    return d_theta.cpu().numpy(), piv, rep


@pytest.mark.parametrize("precision", [isdf.ISDF_PREC_FP32, isdf.ISDF_PREC_BF16X3])
def test_batched_matches_single_solves_and_replays(precision):
    batch = make_batch(31, 6, partial=4)
    inst = to_instances(batch)
    reps, brep = isdf.solve_batched(inst, precision=precision)
    torch.cuda.synchronize()
    assert brep["graph_replayed"] == 0
    first = [a["d_theta"].cpu().numpy() for a in inst]
    for i, a in enumerate(inst):
        th, piv, rep = single(a, precision)
        np.testing.assert_array_equal(a["piv"], piv)
        assert reps[i]["rk_kmeans"] == rep["rk_kmeans"]
        assert reps[i]["kmeans_iters"] == rep["kmeans_iters"]
        assert np.isfinite(first[i]).all()
        err = np.linalg.norm(first[i] - th) / np.linalg.norm(th)
        assert err < 1e-12, (i, err)
    # The partial-density instance went through the single solve.
    assert reps[4]["npt"] < inst[4]["domain"].NumGridTotal()
    for rep_i in range(3):
        for a in inst:  # eager work between replays
            a["d_theta"].fill_(0.0)
            a["piv"][:] = -1
        torch.cuda.synchronize()
        reps2, brep2 = isdf.solve_batched(inst, precision=precision)
        torch.cuda.synchronize()
        assert brep2["graph_replayed"] == 1
        for i, a in enumerate(inst):
            np.testing.assert_array_equal(a["d_theta"].cpu().numpy().view(np.uint64),
                                          first[i].view(np.uint64))
            np.testing.assert_array_equal(a["piv"], single(a, precision)[1])
            assert reps2[i]["kmeans_iters"] == reps[i]["kmeans_iters"]


@pytest.mark.parametrize("precision", [isdf.ISDF_PREC_FP32, isdf.ISDF_PREC_BF16X3])
def test_batched_replay_on_new_data(precision):
    # The captured graph replayed on other wavefunctions written into the same
    # buffers (same shapes): piv and Theta must be those of the new data, so
    # nothing of the previous call (piv staging on the host included) survives.
    batch = make_batch(41, 5)
    inst = to_instances(batch)
    isdf.solve_batched(inst, precision=precision)
    torch.cuda.synchronize()
    for seed in (42, 43):
        rng = np.random.default_rng(seed)
        for a, b in zip(inst, batch):
            psi = synth.cholesky_qr2(rng.standard_normal(b["psi"].shape),
                                     float(np.prod(b["L"])) / b["psi"].shape[1])
            a["d_psi"].copy_(torch.from_numpy(psi))
            a["piv"][:] = -1
        torch.cuda.synchronize()
        _, brep = isdf.solve_batched(inst, precision=precision)
        torch.cuda.synchronize()
        assert brep["graph_replayed"] == 1
        for i, a in enumerate(inst):
            th, piv, _ = single(a, precision)
            np.testing.assert_array_equal(a["piv"], piv, err_msg="seed %d instance %d" % (seed, i))
            err = np.linalg.norm(a["d_theta"].cpu().numpy() - th) / np.linalg.norm(th)
            assert err < 1e-12, (seed, i, err)


def test_batched_pivots_match_oracle():
    batch = make_batch(7, 3)
    inst = to_instances(batch)
    isdf.solve_batched(inst, precision=isdf.ISDF_PREC_BF16X3)
    torch.cuda.synchronize()
    for b, a in zip(batch, inst):
        w = oracle.weights(b["psi"], b["nocc"], b["nocc"], b["nocc"], b["nocc"], b["nocc"])
        km = oracle.kmeans(w, b["mu"], b["iters"], b["L"], b["grid"])
        np.testing.assert_array_equal(a["piv"], km["piv"])


def test_batched_per_instance_iteration_limits():
    # Each instance stops at its own K-means iteration limit (0, 1, 3, 7 and 20
    # in one batch): the loop control of iteration s-1 fused with the update of
    # iteration s, and the last control after the loop, must leave every
    # instance's piv and iteration count as the oracle and its own single
    # solve give them.
    batch = make_batch(29, 5)
    for b, it in zip(batch, (0, 1, 3, 7, 20)):
        b["iters"] = it
    inst = to_instances(batch)
    reps, _ = isdf.solve_batched(inst, precision=isdf.ISDF_PREC_BF16X3)
    torch.cuda.synchronize()
    for i, (b, a, rep) in enumerate(zip(batch, inst, reps)):
        w = oracle.weights(b["psi"], b["nocc"], b["nocc"], b["nocc"], b["nocc"], b["nocc"])
        km = oracle.kmeans(w, b["mu"], b["iters"], b["L"], b["grid"])
        np.testing.assert_array_equal(a["piv"], km["piv"], err_msg="instance %d" % i)
        _, piv, srep = single(a, isdf.ISDF_PREC_BF16X3)
        np.testing.assert_array_equal(piv, a["piv"], err_msg="instance %d" % i)
        assert rep["kmeans_iters"] == srep["kmeans_iters"] == km["iters"], (i, rep, srep)


@pytest.mark.parametrize("shape", ["one", "tiny"])
def test_batched_one_instance_and_tiny_instance(shape):
    # Edge cases of the batch's block lists: a batch of one instance, and an
    # 8 x 8 x 8 instance (two assignment bricks, rk 32) beside a 20^3 one.
    rng = np.random.default_rng(808)
    cases = [((16, 18, 16), 16)] if shape == "one" else [((8, 8, 8), 8), ((20, 20, 20), 16)]
    batch = []
    for grid, nstate in cases:
        L = tuple(0.32 * g for g in grid)
        ng = grid[0] * grid[1] * grid[2]
        psi = synth.cholesky_qr2(rng.standard_normal((nstate, ng)), L[0] * L[1] * L[2] / ng)
        batch.append(dict(psi=psi, grid=grid, L=L, nstate=nstate, nocc=nstate // 2,
                          mu=8 * (nstate // 2), iters=12))
    inst = to_instances(batch)
    isdf.solve_batched(inst, precision=isdf.ISDF_PREC_BF16X3)
    torch.cuda.synchronize()
    for i, (b, a) in enumerate(zip(batch, inst)):
        w = oracle.weights(b["psi"], b["nocc"], b["nocc"], b["nocc"], b["nocc"], b["nocc"])
        km = oracle.kmeans(w, b["mu"], b["iters"], b["L"], b["grid"])
        np.testing.assert_array_equal(a["piv"], km["piv"], err_msg="instance %d" % i)
        th, piv, _ = single(a, isdf.ISDF_PREC_BF16X3)
        np.testing.assert_array_equal(piv, a["piv"], err_msg="instance %d" % i)
        th0 = a["d_theta"].cpu().numpy()
        assert np.isfinite(th0).all()
        err = np.linalg.norm(th0 - th) / np.linalg.norm(th)
        assert err < 1e-12, (i, err)


def test_batched_ragged_rk_above_one_block():
    # rk 136, 264 and 384 (nstate 34, 66, 96): several 128-wide block columns
    # per instance, so the off-diagonal F / G blocks, the forward-only identity
    # solve across block columns and the batched X X^T on instances of
    # different n all run.
    rng = np.random.default_rng(13)
    batch = []
    for i, (n, nstate) in enumerate(((24, 34), (28, 66), (32, 96))):
        grid = (n, n, n - 4 * (i % 2))
        L = tuple(0.32 * g for g in grid)
        ng = grid[0] * grid[1] * grid[2]
        psi = synth.cholesky_qr2(rng.standard_normal((nstate, ng)), L[0] * L[1] * L[2] / ng)
        batch.append(dict(psi=psi, grid=grid, L=L, nstate=nstate, nocc=nstate // 2,
                          mu=8 * (nstate // 2), iters=15))
    assert [b["mu"] for b in batch] == [136, 264, 384]
    for precision in (isdf.ISDF_PREC_FP32, isdf.ISDF_PREC_BF16X3):
        inst = to_instances(batch)
        reps, brep = isdf.solve_batched(inst, precision=precision)
        torch.cuda.synchronize()
        first = [a["d_theta"].cpu().numpy() for a in inst]
        for b, a, th0 in zip(batch, inst, first):
            w = oracle.weights(b["psi"], b["nocc"], b["nocc"], b["nocc"], b["nocc"], b["nocc"])
            km = oracle.kmeans(w, b["mu"], b["iters"], b["L"], b["grid"])
            np.testing.assert_array_equal(a["piv"], km["piv"])
            th, piv, rep = single(a, precision)
            np.testing.assert_array_equal(piv, km["piv"])
            assert np.isfinite(th0).all()
            err = np.linalg.norm(th0 - th) / np.linalg.norm(th)
            assert err < 1e-12, err
        for _ in range(2):
            for a in inst:
                a["d_theta"].fill_(0.0)
            torch.cuda.synchronize()
            _, brep2 = isdf.solve_batched(inst, precision=precision)
            torch.cuda.synchronize()
            assert brep2["graph_replayed"] == 1
            for a, th0 in zip(inst, first):
                np.testing.assert_array_equal(a["d_theta"].cpu().numpy().view(np.uint64),
                                              th0.view(np.uint64))
