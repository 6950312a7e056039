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
    # the C++ drop-in entry point, isdf::isdffunc::ISDF(double*, Domain, ...)
    assert "_ZN4isdf8isdffunc4ISDFEPdNS_6DomainEiiiiiiiS1_Piib" in out


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
])
def test_invalid_arguments_rejected_before_device_work(kw, msg):
    dom = isdf.Domain((10.0,) * 3, (16, 16, 16))
    a = dict(nocc=8, nstate=16, nv1=8, nc1=4, nv2=8, nc2=4, mu_points=64, max_iter=5,
             precision=isdf.ISDF_PREC_FP64)
    a.update(kw)
    piv = np.zeros(4096, np.int32)
    with pytest.raises(ValueError, match=msg):
        isdf.solve(1, dom, a["nocc"], a["nstate"], a["nv1"], a["nc1"], a["nv2"], a["nc2"],
                   a["mu_points"], 1, piv, a["max_iter"], stream=0, precision=a["precision"])


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
