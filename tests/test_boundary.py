"""The C ABI boundary (include/isdf_hip.h) and the host-side code around it,
checked without a GPU: the library loads, exports every declared symbol, the
ctypes structs match the C header's layout, arguments are validated before any
device work, and the check=true validator reproduces the reference's numbers."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import libisdf_amd as isdf
import oracle
from libisdf_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "isdf_hip.h")).read()
    return sorted(set(re.findall(r"\b(isdf_hip_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = header_functions()
    assert set(declared) == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True).stdout
    for name in declared:
        assert re.search(r"\bT %s$" % name, out, re.M), name
    # the C++ drop-in entry points, isdf::isdffunc::ISDF(double*, Domain, ...): the
    # device-pointer one (..., bool check) and the host-pointer one (..., int
    # scalblocksize, std::string s, std::string s1, bool check)
    assert "_ZN4isdf8isdffunc4ISDFEPdNS_6DomainEiiiiiiiS1_Piib" in out
    assert ("_ZN4isdf8isdffunc4ISDFEPdNS_6DomainEiiiiiiiS1_PiiiNSt7__cxx1112basic_stringIcSt11"
            "char_traitsIcESaIcEEES9_b") in out


def test_struct_layout_matches_c_header(tmp_path):
    prog = tmp_path / "sz.c"
    prog.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "isdf_hip.h"\n'
                    'int main(void){printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(isdf_problem),'
                    ' sizeof(isdf_report), sizeof(isdf_debug), sizeof(isdf_check),'
                    ' offsetof(isdf_problem, ld_psi), offsetof(isdf_report, ms_total));'
                    'printf("%zu %zu\\n", sizeof(isdf_batch_report),'
                    ' offsetof(isdf_batch_report, theta_tile_ms));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)])
    got = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    want = [ctypes.sizeof(_lib.IsdfProblem), ctypes.sizeof(_lib.IsdfReport),
            ctypes.sizeof(_lib.IsdfDebug), ctypes.sizeof(_lib.IsdfCheck),
            _lib.IsdfProblem.ld_psi.offset, _lib.IsdfReport.ms_total.offset,
            ctypes.sizeof(_lib.IsdfBatchReport), _lib.IsdfBatchReport.theta_tile_ms.offset]
    assert got == want


def test_rank_rule_matches_oracle():
    for args in [(16, 0, 16, 0, 0), (3, 0, 5, 0, 0), (2, 1, 1, 0, 0), (64, 64, 64, 64, 1024)]:
        assert isdf.isdf_rank(*args) == oracle.irank(args[4], *args[:4])


@pytest.mark.parametrize("kw,msg", [
    (dict(mu_points=5000), "mu_points exceeds"),
    (dict(nv1=9), "pair ranges"),
    (dict(nc2=20), "pair ranges"),
    (dict(max_iter=-1), "kmeans_max_iter"),
    (dict(precision=7), "unknown precision"),
    (dict(precision=isdf.ISDF_PREC_REFINE), "mixed forms only"),
    (dict(select=2), "unknown point selection"),
])
def test_invalid_arguments_rejected_before_device_work(kw, msg):
    dom = isdf.Domain((10.0,) * 3, (16, 16, 16))
    a = dict(nocc=8, nstate=16, nv1=8, nc1=4, nv2=8, nc2=4, mu_points=64, max_iter=5,
             precision=isdf.ISDF_PREC_FP64, select=0)
    a.update(kw)
    piv = np.zeros(4096, np.int32)
    with pytest.raises(ValueError, match=msg):
        isdf.solve(1, dom, a["nocc"], a["nstate"], a["nv1"], a["nc1"], a["nv2"], a["nc2"],
                   a["mu_points"], 1, piv, a["max_iter"], stream=0, precision=a["precision"],
                   select=a.get("select", 0))


def test_check_validator_reproduces_reference_si8(si8):
    psi, ref = si8
    a = ref["args"]
    r = oracle.isdf(psi, (10.216,) * 3, (16, 16, 16), a["nocc"], a["nstate"], a["nv1"], a["nc1"],
                    a["nv2"], a["nc2"], a["mu_points"], a["kmeans_max_iter"])
    dom = isdf.Domain((10.216,) * 3, (16, 16, 16))
    c = isdf.check(psi, r["theta"], r["piv"], dom, a["nocc"], a["nstate"], a["nv1"], a["nc1"],
                   a["nv2"], a["nc2"], a["mu_points"])
    assert float("%.6g" % c["fit_error"]) == ref["fit_error"]
    assert float("%.6g" % c["coulomb_error"]) == ref["coulomb_error"]
    # and against the oracle's restatements beyond the printed digits
    fit = oracle.fit_error(psi, a["nocc"], a["nv1"], a["nc1"], a["nv2"], a["nc2"], r["theta"],
                           r["piv"])
    cou = oracle.coulomb_error(psi, (10.216,) * 3, (16, 16, 16), a["nocc"], a["nv1"], a["nc1"],
                               a["nv2"], a["nc2"], r["theta"], r["piv"])
    assert c["fit_error"] == pytest.approx(fit, rel=1e-10)
    assert c["coulomb_error"] == pytest.approx(cou, rel=1e-8)


def test_check_validator_dual_pair_sets():
    rng = np.random.default_rng(12)
    n = (5, 6, 4)
    psi = rng.standard_normal((10, 120))
    theta = rng.standard_normal((30, 120))
    piv = rng.permutation(120).astype(np.int32)
    dom = isdf.Domain((3.0, 3.5, 2.0), n)
    c = isdf.check(psi, theta, piv, dom, 5, 10, 4, 3, 2, 5, 30)
    fit = oracle.fit_error(psi, 5, 4, 3, 2, 5, theta, piv)
    cou = oracle.coulomb_error(psi, (3.0, 3.5, 2.0), n, 5, 4, 3, 2, 5, theta, piv)
    assert c["fit_error"] == pytest.approx(fit, rel=1e-12)
    assert c["coulomb_error"] == pytest.approx(cou, rel=1e-9)


@pytest.mark.parametrize("prec,msg", [
    (isdf.ISDF_PREC_FP64, "mixed forms"),
    (isdf.ISDF_PREC_FP32 | isdf.ISDF_PREC_REFINE, "REFINE is not supported"),
])
def test_batched_rejects_before_device_work(prec, msg):
    dom = isdf.Domain((10.0,) * 3, (16, 16, 16))
    inst = [dict(d_psi=1, domain=dom, nocc=8, nstate=16, nv1=8, nc1=8, nv2=8, nc2=8,
                 mu_points=64, d_theta=1, piv=np.zeros(4096, np.int32), max_iter=5)]
    with pytest.raises(ValueError, match=msg):
        isdf.solve_batched(inst, precision=prec, stream=0)
    inst[0]["nv2"] = 4  # differing pair sets
    with pytest.raises(ValueError, match="pair sets|mixed forms"):
        isdf.solve_batched(inst, precision=isdf.ISDF_PREC_FP32, stream=0)


def test_solve_host_rejects_before_device_work():
    L = _lib.lib()
    dom = isdf.Domain((10.0,) * 3, (8, 8, 8))
    p = isdf._problem(dom, 8, 16, 8, 4, 8, 4, 64, 5, isdf.ISDF_PREC_FP64)
    buf = np.zeros(512 * 64, np.float64)
    piv = np.zeros(512, np.int32)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.isdf_hip_solve_host(ctypes.byref(p), None, buf.ctypes.data_as(dp),
                               piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None, None, None)
    assert rc == isdf.ISDF_EINVAL and "NULL buffer" in _lib.last_error()
    p.nv1 = 9
    rc = L.isdf_hip_solve_host(ctypes.byref(p), buf.ctypes.data_as(dp), buf.ctypes.data_as(dp),
                               piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None, None, None)
    assert rc == isdf.ISDF_EINVAL and "pair ranges" in _lib.last_error()
    with pytest.raises(ValueError, match="theta must be"):
        isdf.solve_host(np.zeros((16, 512)), dom, 8, 16, 8, 4, 8, 4, 64, np.zeros(10), piv, 5)


def test_host_interface_dispatch_on_s_and_s1():
    """isdf.cpp:44, :60-74, :98-113: piv is zeroed first; s1 other than Kmeans
    or QRCP is "Wrong select points method"; an unknown s computes nothing."""
    dom = isdf.Domain((10.0,) * 3, (8, 8, 8))
    psi = np.zeros((16, 512))
    theta = np.full((64, 512), 3.0)
    piv = np.full(512, 5, np.int32)
    assert isdf.ISDF_host(psi, dom, 8, 16, 8, 4, 8, 4, 64, theta, piv, 5, s="other") == {}
    assert (piv == 0).all() and (theta == 3.0).all()
    with pytest.raises(RuntimeError, match="Wrong select points method"):
        isdf.ISDF_host(psi, dom, 8, 16, 8, 4, 8, 4, 64, theta, piv, 5, s1="Random")
    with pytest.raises(RuntimeError, match="Wrong select points method"):
        isdf.ISDF_host(psi, dom, 8, 16, 8, 4, 8, 4, 64, theta, piv, 5, s="MPI", s1="qrcp")


CALLER_HOST = os.path.join(ROOT, "tests", "dropin", "_build", "caller_host")


@pytest.mark.parametrize("s,s1,status,msg", [
    ("other", "Kmeans", 0, None),
    ("scalapack", "Random", 2, "Wrong select points method"),
    ("MPI", "qrcp", 2, "Wrong select points method"),
])
def test_cpp_host_interface_dispatch(tmp_path, s, s1, status, msg):
    """The C++ host-pointer ISDF (built from include/isdf.hpp without -DGPU) on
    the paths that end before any device work: ErrorHandling's log lines and
    exception, and the zeroed piv of an unknown s."""
    ng, nstate, rk = 512, 16, 64
    np.zeros(ng * nstate).tofile(tmp_path / "psi.bin")
    args = [CALLER_HOST, str(tmp_path / "psi.bin"), "8", "8", "8", "10", "10", "10", "8",
            str(nstate), "8", "4", "8", "4", str(rk), "5", "0", str(tmp_path / "theta.bin"),
            str(tmp_path / "piv.bin"), str(tmp_path / "isdf.0"), s, s1, "32"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert r.returncode == status, r.stderr
    log = (tmp_path / "isdf.0").read_text()
    assert log.startswith("rk=%d\n" % rk)
    assert (np.fromfile(tmp_path / "piv.bin", np.int32) == 0).all()
    assert (np.fromfile(tmp_path / "theta.bin") == 0.0).all()
    if msg:
        assert "ERROR!" in log and msg in log and msg in r.stderr


def test_fit_error_rejects_before_device_work():
    dom = isdf.Domain((10.0,) * 3, (8, 8, 8))
    piv = np.zeros(512, np.int32)
    piv[3] = 512
    with pytest.raises(ValueError, match=r"piv\[3\] is not a grid index"):
        isdf.fit_error(1, 1, piv, dom, 8, 16, 8, 4, 8, 4, 64, stream=0)
    with pytest.raises(ValueError, match="pair ranges"):
        isdf.fit_error(1, 1, np.zeros(512, np.int32), dom, 8, 16, 9, 4, 8, 4, 64, stream=0)


def test_device_buffer_size_rule():
    # psi: column-major Ng x nstate with leading dimension ld spans
    # ld (nstate - 1) + Ng doubles; Theta: Ng x rk.
    assert isdf.buffer_sizes(1000, 8, 64) == (8000, 64000)
    assert isdf.buffer_sizes(1000, 8, 64, ld_psi=1024) == (1024 * 7 + 1000, 64000)
    assert isdf.buffer_sizes(10, 1, 3) == (10, 30)
    with pytest.raises(ValueError, match="below Ng"):
        isdf.buffer_sizes(1000, 8, 64, ld_psi=999)


def test_device_entry_points_reject_host_and_short_buffers():
    # No device work happens: every argument error is raised in Python first.
    torch = pytest.importorskip("torch")
    dom = isdf.Domain((4.0, 4.0, 4.0), (8, 8, 8))
    piv = np.zeros(512, np.int32)
    host = torch.zeros(16 * 512, dtype=torch.float64)
    with pytest.raises(ValueError, match="device buffer expected"):
        isdf.solve(host, dom, 8, 16, 8, 8, 8, 8, 64, host, piv, 10)
    with pytest.raises(ValueError, match="at least 512"):
        isdf.solve(host, dom, 8, 16, 8, 8, 8, 8, 64, host, piv[:100], 10)
    with pytest.raises(ValueError, match="rk = 64 are needed"):
        isdf.fit_error(host, host, piv[:10], dom, 8, 16, 8, 8, 8, 8, 64)
