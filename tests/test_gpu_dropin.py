"""The C++ drop-ins of include/isdf.hpp, called the way LibISDF's own example
calls them (examples/si8/readpsi.cpp:170-240), on the reference's Si8 case:
the device-pointer ISDF (built with -DGPU, include/isdf.hpp:37) and the
host-pointer one (include/isdf.hpp:39), both with check=true.  Their isdfOFS
logs must carry the reference's printed K-means trace and fit error, and
their Theta and piv must equal the C-ABI solve's bit for bit (the same
kernels on the same input).  Also the Python host mirror, solve_host."""
import os
import re
import subprocess

import numpy as np
import pytest

import libisdf_amd as isdf

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "dropin", "_build")
L, G = (10.216,) * 3, (16, 16, 16)


def device_solve(psi, a):
    rk = a["mu_points"]
    d_psi = torch.from_numpy(np.ascontiguousarray(psi)).cuda()
    d_theta = torch.empty((rk, psi.shape[1]), dtype=torch.float64, device="cuda")
    piv = np.zeros(psi.shape[1], np.int32)
    isdf.solve(d_psi, isdf.Domain(L, G), a["nocc"], a["nstate"], a["nv1"], a["nc1"], a["nv2"],
               a["nc2"], rk, d_theta, piv, a["kmeans_max_iter"])
    torch.cuda.synchronize()
    return d_theta.cpu().numpy(), piv


def run_caller(tmp_path, which, psi, a):
    ng = psi.shape[1]
    f = lambda n: str(tmp_path / ("%s_%s" % (which, n)))  # noqa: E731
    np.ascontiguousarray(psi).tofile(f("psi.bin"))
    args = [os.path.join(BUILD, "caller_" + which), f("psi.bin"), *map(str, G), *map(str, L),
            *(str(a[k]) for k in ("nocc", "nstate", "nv1", "nc1", "nv2", "nc2", "mu_points",
                                  "kmeans_max_iter")), "1", f("theta.bin"), f("piv.bin"),
            f("isdf.0")]
    print("drop-in caller", which, flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    theta = np.fromfile(f("theta.bin")).reshape(a["mu_points"], ng)
    piv = np.fromfile(f("piv.bin"), np.int32)
    return theta, piv, open(f("isdf.0")).read()


@pytest.mark.parametrize("which", ["gpu", "host"])
def test_cpp_dropin_si8(tmp_path, si8, which):
    psi, ref = si8
    a = ref["args"]
    theta, piv, log = run_caller(tmp_path, which, psi, a)
    want_theta, want_piv = device_solve(psi, a)
    assert np.array_equal(piv, want_piv)
    assert np.array_equal(theta, want_theta)
    lines = log.splitlines()
    assert lines[0] == "rk=%d" % a["mu_points"]
    assert lines[1].split() == [str(t) for t in ref["kmeans_trace"]]
    assert lines[2] == "Converged in %d iterations." % ref["kmeans_iters"]
    fit = re.search(r"sqrt\(\(rhoij-rho\^hat\)\^2\)/sqrt\(rhoij\^2\)= (\S+)", log).group(1)
    assert float(fit) == ref["fit_error"]
    cou = re.search(r"sqrt\(\(ijkl-ijkl\^hat\)\^2\)/sqrt\(ijkl\^2\) =(\S+)", log).group(1)
    # Theta is ill-conditioned here (kappa ~ 2e11), so the Coulomb error agrees
    # with the reference's printed one to ~1e-4 relative.
    assert float(cou) == pytest.approx(ref["coulomb_error"], rel=1e-4)


def test_python_solve_host_equals_device_solve(si8):
    psi, ref = si8
    a = ref["args"]
    rk, ng = a["mu_points"], psi.shape[1]
    theta = np.full((rk, ng), np.nan)
    piv = np.full(ng, -3, np.int32)
    trace = np.zeros(a["kmeans_max_iter"], np.int32)
    rep = isdf.solve_host(np.ascontiguousarray(psi), isdf.Domain(L, G), a["nocc"], a["nstate"],
                          a["nv1"], a["nc1"], a["nv2"], a["nc2"], rk, theta, piv,
                          a["kmeans_max_iter"], debug=dict(trace=trace))
    assert trace[:rep["kmeans_iters"]].tolist() == ref["kmeans_trace"]
    want_theta, want_piv = device_solve(psi, a)
    assert np.array_equal(piv, want_piv) and np.array_equal(theta, want_theta)
    # a second call on a larger problem regrows the cached device copies
    psi2 = np.vstack([psi, psi[:5]])
    theta2 = np.empty((rk, ng))
    isdf.solve_host(psi2, isdf.Domain(L, G), a["nocc"], a["nstate"] + 5, a["nv1"], a["nc1"],
                    a["nv2"], a["nc2"], rk, theta2, piv, a["kmeans_max_iter"])
    assert np.array_equal(theta2, want_theta)
