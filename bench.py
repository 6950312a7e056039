"""Benchmark of the ISDF hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cfg3]

One step is one full ISDF solve (weights -> K-means -> piv -> Psi(mu) ->
Z C^T -> C C^T + eps -> potrf -> Theta) through the C ABI, with psi already in
HBM and piv written back to the host, as the reference's ISDF() call does.
The default workload is the 96^3 configuration the mixed-precision targets are
stated on (Ng = 884736, nstate = 256, all pair counts 128, rk = 2048, 50
K-means iterations) in FP64 on seeded synthetic wavefunctions; the BF16X3,
FP32 and FP32 + FP64-correction forms are then timed on the same input and
compared with the FP64 result.  Prints one JSON line.

`roofline` covers the dominant kernel (the in-place blocked triangular solve
that produces Theta): algorithmic flops 2 Ng rk^2 per launch over its average
duration, timed with HIP events recorded on the solve's own stream.
`cpu_baseline` times the CPU oracle on a bounded sample of the same workload
on this host (BLAS-3 for the GEMM, Cholesky and triangular-solve stages, as
the reference's MKL calls) and extrapolates to one solve.
`--peak-probe` runs only register-only loops of the FP64 / FP32 units.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ISDF solves/s and ZC^T TFLOP/s, mixed vs FP64 plain at fixed fit error, 1 MI355X"
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3     # dense FP32 MFMA
BF16_PEAK_TFLOPS = 2500.0    # dense bf16 MFMA
HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the triangular solve (k_potrs_dpp, the default) from two rocprofv3
# counter passes (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000, factors from a copy kernel of known
# size), per workload.
TRAFFIC = {
    "cfg3": {"bytes": 2.9289e11,
             "source": "profiles/k_potrs_k_zct_cfg3_fp64_hbm_bytes.md (263.90 GB read + "
                       "28.99 GB written per launch; 57.98 GB algorithmic)"},
    "cfg2": {"bytes": 2.422e10,
             "source": "profiles/k_potrs_k_zct_cfg2_fp64_hbm_bytes.md (19.93 GB read + "
                       "4.295 GB written per launch; 8.59 GB algorithmic)"},
}


def flops(w, rk):
    k1 = w.nv1 + w.nc1
    k2 = w.nv2 + w.nc2
    dual = (w.nv1 != w.nv2 or w.nc1 != w.nc2)
    zct = 2.0 * w.ng * rk * k1 + (2.0 * w.ng * rk * k2 if dual else 0.0) + w.ng * rk
    potrs = 2.0 * w.ng * rk * rk
    potrf = rk ** 3 / 3.0
    fit = 2.0 * w.ng * rk * k1 * k2  # residual GEMM Theta (Ng x rk) x C (rk x K1K2)
    return dict(zct=zct, potrs=potrs, potrf=potrf, fit=fit)


def cpu_baseline(psi, w, iters):
    """Time the CPU oracle on a bounded sample of the workload and extrapolate
    one solve: weights (all rows), K-means (setup + 2 Lloyd iterations, scaled
    to the GPU's iteration count), Z C^T (dgemm + Hadamard) and the two
    triangular solves (dtrsm) on a row sample scaled to Ng, and the full
    rk x rk Cholesky (dpotrf).  The BLAS/LAPACK stages use the threaded BLAS
    numpy links, as the reference uses MKL."""
    import oracle
    rk = w.mu_points
    t = {}
    t0 = time.perf_counter()
    wts = oracle.weights(psi, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2)
    t["weights"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    km0 = oracle.kmeans(wts, rk, 0, w.length, w.ngrid)
    ta = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.kmeans(wts, rk, 2, w.length, w.ngrid)
    tb = time.perf_counter() - t0
    per_it = max(tb - ta, 0.0) / 2.0
    t["kmeans"] = ta + per_it * iters
    piv = km0["piv"]
    psimu = oracle.gather(psi, piv, rk)
    nrow = min(w.ng, 262144)
    rows = np.arange(nrow, dtype=np.int64)
    t0 = time.perf_counter()
    Z = oracle.zct_rows_blas(psi, psimu, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2, rows)
    t["zct"] = (time.perf_counter() - t0) * w.ng / nrow
    P = oracle.zct_rows_blas(psi, psimu, w.nocc, w.nv1, w.nc1, w.nv2, w.nc2,
                             piv[:rk].astype(np.int64))
    P[np.diag_indices(rk)] += 1e-10 * np.trace(P) / rk
    t0 = time.perf_counter()
    Lc = oracle.potrf_lapack(P)
    t["potrf"] = time.perf_counter() - t0
    nsol = min(nrow, 65536)
    t0 = time.perf_counter()
    oracle.potrs_rows_blas(Lc, Z[:nsol])
    t["potrs"] = (time.perf_counter() - t0) * w.ng / nsol
    total = sum(t.values())
    omp = int(oracle.lib().orc_num_threads())
    blas = oracle.blas_threads()
    return dict(value=1.0 / total, unit="solves/s", cores=max(omp, blas or 0),
                kind="port",
                sample=("CPU oracle on the same %s input (%d^3, rk %d): weights over all rows "
                        "(OpenMP, %d threads); K-means setup + 2 Lloyd iterations scaled to %d "
                        "(4-wide vector distance loops, OpenMP); Z C^T as dgemm + Hadamard on %d "
                        "rows and the two triangular solves as dtrsm on %d rows, scaled to Ng; "
                        "the full rk x rk dpotrf (OpenBLAS, %s threads); est. %.1f s per solve"
                        % (w.name, w.n, rk, omp, iters, nrow, nsol, blas, total)),
                stage_s={k: round(v, 4) for k, v in t.items()})


def peak_probe():
    """Register-only loops of v_mfma_f64_16x16x4f64, v_fma_f64 and
    v_mfma_f32_16x16x4f32 over the whole chip (libisdf_probe.so)."""
    import ctypes
    import torch  # noqa: F401  (initialises the HIP runtime in this process)
    lib = ctypes.CDLL(os.path.join(ROOT, "libisdf_amd", "libisdf_probe.so"))
    lib.isdf_probe_peak.restype = ctypes.c_int
    lib.isdf_probe_peak.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    out = {}
    for which, name, peak, iters in ((0, "v_mfma_f64_16x16x4f64", FP64_PEAK_TFLOPS, 4000),
                                     (1, "v_fma_f64", FP64_PEAK_TFLOPS, 20000),
                                     (2, "v_mfma_f32_16x16x4f32", FP32_PEAK_TFLOPS, 8000),
                                     (5, "v_fmac_f64_dpp row_newbcast (64 accumulators)",
                                      FP64_PEAK_TFLOPS, 5000)):
        res = []
        # the DPP loop holds 138 VGPRs (3 waves per SIMD at most): also 1, 2, 3 waves
        for blocks in ((256, 512, 768, 256 * 4) if which == 5 else (256 * 4, 256 * 8)):
            tf, ms = ctypes.c_double(), ctypes.c_double()
            rc = lib.isdf_probe_peak(which, blocks, iters, ctypes.byref(tf), ctypes.byref(ms))
            if rc != 0:
                raise RuntimeError("probe %s failed: hipError %d" % (name, rc))
            res.append({"workgroups": blocks, "waves_per_simd": blocks // 256,
                        "tflops": tf.value, "ms": ms.value, "frac_of_peak": tf.value / peak})
        out[name] = {"published_peak_tflops": peak, "runs": res}
    # Both FP64 pipes at once: half the waves on MFMA, half on vector FMA.
    res = []
    for blocks in (256 * 4, 256 * 8):
        tf, ms = ctypes.c_double(), ctypes.c_double()
        rc = lib.isdf_probe_peak(4, blocks, 2000, ctypes.byref(tf), ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError("mixed-pipe probe failed: hipError %d" % rc)
        res.append({"workgroups": blocks, "tflops": tf.value, "ms": ms.value,
                    "frac_of_peak": tf.value / FP64_PEAK_TFLOPS})
    out["fp64_mfma_and_fma_waves_together"] = {"published_peak_tflops": FP64_PEAK_TFLOPS,
                                               "runs": res}
    # Latency of back-to-back dependent FP64 adds (the floor of a strictly
    # sequential sum such as the K-means centroid sums): 256 single-wave groups.
    tf, ms = ctypes.c_double(), ctypes.c_double()
    rc = lib.isdf_probe_peak(3, 256, 200000, ctypes.byref(tf), ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError("add-chain probe failed: hipError %d" % rc)
    out["v_add_f64_dependent_chain"] = {"ns_per_add": 1e3 / (tf.value * 1e6), "ms": ms.value}
    print(json.dumps({"peak_probe": out}))


def compare_solve_tiles(a):
    """FP64 solves of one workload with each listed triangular-solve tile (and Z C^T
    tile), interleaved in one process so that every tile sees the same clock and
    thermal history."""
    import torch
    import libisdf_amd as isdf
    from libisdf_amd import synth
    w = synth.WORKLOADS[a.workload]
    psi = synth.wavefunctions(w)
    rk = isdf.isdf_rank(w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points)
    fl = flops(w, rk)
    d_psi = torch.from_numpy(psi).cuda()
    d_theta = torch.empty((rk, w.ng), dtype=torch.float64, device="cuda")
    first = None
    names = ("ISDF_POTRS_TILE", "ISDF_ZCT_TILE")
    saved = {k: os.environ.get(k) for k in names}
    try:
        for entry in a.compare_solve_tiles.split(","):
            parts = entry.split("/")
            for k, v in zip(names, parts + [None]):
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
            piv = np.zeros(w.ng, np.int32)
            rep = isdf.solve(d_psi, isdf.Domain(w.length, w.ngrid), w.nocc, w.nstate, w.nv1,
                             w.nc1, w.nv2, w.nc2, w.mu_points, d_theta, piv, w.kmeans_iters)
            torch.cuda.synchronize()
            if first is None:
                first = d_theta.clone()
                same = True
            else:
                same = bool(torch.equal(d_theta, first))
            print(json.dumps({"workload": w.name, "tiles": entry, "potrs_ms": rep["ms_potrs"],
                              "potrs_tflops": fl["potrs"] / rep["ms_potrs"] / 1e9,
                              "zct_ms": rep["ms_zct"], "zct_tflops": fl["zct"] / rep["ms_zct"] / 1e9,
                              "total_ms": rep["ms_total"], "theta_bitwise_equal_first": same,
                              "stage_ms": {k[3:]: round(v, 3) for k, v in rep.items()
                                           if k.startswith("ms_")}}),
                  flush=True)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def meets_target(m, fit64):
    """The mixed-form target of the 96^3 configuration: >= 3x the FP64 solve,
    Theta within 1e-6 of FP64, fit error within 1.1x of FP64's."""
    return bool(m["speedup_vs_fp64"] >= 3.0 and m["theta_rel_err_vs_fp64"] <= 1e-6
                and m["fit_error"] <= 1.1 * fit64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cfg3")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", dest="mixed", action="store_false",
                    help="skip the mixed-precision solves (FP32 and 3-way-split BF16)")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                    help="run only a plain device copy of GIB GiB five times (the known-size "
                         "kernel that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--check", action="store_true",
                    help="instead of the solve benchmark: time check=true on the device (fit and "
                         "Coulomb errors) against the host validator on the workload's FP64 "
                         "result (small workloads: the Coulomb half costs ~4 (K1 K2)^2 Ng flops)")
    ap.add_argument("--profile-solve", choices=("fp64", "fp32", "bf16x3"), default=None,
                    help="instead of the benchmark: one solve of the workload in this precision "
                         "and nothing else (a fixed program for rocprofv3 counter passes)")
    ap.add_argument("--select", choices=("kmeans", "qrcp"), default="kmeans",
                    help="interpolation-point selection for --profile-solve")
    ap.add_argument("--repeat", type=int, default=1, metavar="N",
                    help="with --profile-solve: run the solve N times back to back (a sustained "
                         "load for sampling board power and clocks)")
    ap.add_argument("--peak-probe", action="store_true",
                    help="run only register-only throughput loops of the FP64 / FP32 units")
    ap.add_argument("--compare-solve-tiles", default=None, metavar="LIST",
                    help="instead of the benchmark: FP64 solves of the workload, one per entry "
                         "of the comma-separated list, in that order in one process; an entry is "
                         "a triangular-solve tile (ISDF_POTRS_TILE) or solve/Z C^T tiles "
                         "('9/5': ISDF_ZCT_TILE too); prints each solve's times and whether "
                         "Theta equals the first solve's bit for bit")
    a = ap.parse_args()
    if a.peak_probe:
        return peak_probe()
    if a.copy_probe > 0:
        import torch
        n = int(a.copy_probe * (1 << 30)) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.add(src, 0.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 8 * n, "copy_probe_bytes_written": 8 * n,
                          "launches": 5}))
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: the ISDF path runs as a single rank with no "
              "collectives; --gpus %d is not supported." % a.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    import libisdf_amd as isdf
    from libisdf_amd import synth

    if a.workload == "cfg5":
        return bench_batch(a)
    if a.check:
        return bench_check(a)
    if a.compare_solve_tiles:
        return compare_solve_tiles(a)
    if a.profile_solve:
        w = synth.WORKLOADS[a.workload]
        psi = synth.wavefunctions(w)
        rk = isdf.isdf_rank(w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points)
        prec = {"fp64": isdf.ISDF_PREC_FP64, "fp32": isdf.ISDF_PREC_FP32,
                "bf16x3": isdf.ISDF_PREC_BF16X3}[a.profile_solve]
        d_psi = torch.from_numpy(psi).cuda()
        d_theta = torch.empty((rk, w.ng), dtype=torch.float64, device="cuda")
        piv = np.zeros(w.ng, np.int32)
        sel = isdf.ISDF_SELECT_QRCP if a.select == "qrcp" else isdf.ISDF_SELECT_KMEANS
        t0 = time.perf_counter()
        for r in range(max(1, a.repeat)):
            rep = isdf.solve(d_psi, isdf.Domain(w.length, w.ngrid), w.nocc, w.nstate, w.nv1,
                             w.nc1, w.nv2, w.nc2, w.mu_points, d_theta, piv, w.kmeans_iters,
                             precision=prec, select=sel)
            if a.repeat > 1 and (r + 1) % 20 == 0:
                print("solve %d of %d" % (r + 1, a.repeat), file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / max(1, a.repeat)
        print(json.dumps({"profile_solve": a.profile_solve, "select": a.select,
                          "workload": w.name, "repeat": max(1, a.repeat), "ms_per_solve": dt * 1e3,
                          "stage_ms": {k[3:]: v for k, v in rep.items() if k.startswith("ms_")}}))
        return
    w = synth.WORKLOADS[a.workload]
    psi = synth.wavefunctions(w)
    rk = isdf.isdf_rank(w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points)
    dev = torch.device("cuda:0")
    d_psi = torch.from_numpy(psi).to(dev)
    dom = isdf.Domain(w.length, w.ngrid)
    stream = torch.cuda.current_stream()
    fl = flops(w, rk)

    def timed(precision, d_theta, piv, steps, warmup):
        def step():
            return isdf.solve(d_psi, dom, w.nocc, w.nstate, w.nv1, w.nc1, w.nv2, w.nc2,
                              w.mu_points, d_theta, piv, w.kmeans_iters, stream=stream,
                              precision=precision)
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        reps = []
        t0 = time.perf_counter()
        for _ in range(steps):
            reps.append(step())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        avg = {k: float(np.mean([r[k] for r in reps])) for k in reps[0] if k.startswith("ms_")}
        return dt, avg, reps[-1]

    # The drop-in FP64 path: the timed value.
    d_theta = torch.empty((rk, w.ng), dtype=torch.float64, device=dev)
    piv = np.zeros(w.ng, np.int32)
    dt, avg, r0 = timed(isdf.ISDF_PREC_FP64, d_theta, piv, a.steps, a.warmup)
    ms_step = dt * 1e3 / a.steps
    potrs_tf = fl["potrs"] / (avg["ms_potrs"] * 1e-3) / 1e12
    zct_tf = fl["zct"] / (avg["ms_zct"] * 1e-3) / 1e12
    traffic = TRAFFIC.get(a.workload)
    # The same FP64 solve with no CUs left to the concurrent Cholesky branch:
    # the ZC^T kernel then holds the whole chip while it runs (its standalone
    # rate); the default solve trades some ZC^T rate for the overlap.
    saved = os.environ.get("ISDF_ZCT_RESERVE_CUS")
    os.environ["ISDF_ZCT_RESERVE_CUS"] = "0"
    try:
        _, avg_alone, _ = timed(isdf.ISDF_PREC_FP64, d_theta, piv, 2, 1)
    finally:
        if saved is None:
            del os.environ["ISDF_ZCT_RESERVE_CUS"]
        else:
            os.environ["ISDF_ZCT_RESERVE_CUS"] = saved
    zct_alone_tf = fl["zct"] / (avg_alone["ms_zct"] * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": a.steps / dt,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded cosine plane waves, FP64 CholeskyQR2-orthonormalised)",
        "config": {
            "workload": "%s: %d^3 grid (Ng=%d), L=%.2f bohr, nstate=%d, nocc=%d, "
                        "nv1=nc1=%d, nv2=nc2=%d, mu_points=%d, %d K-means iterations, FP64 plain"
                        % (w.name, w.n, w.ng, w.L, w.nstate, w.nocc, w.nv1, w.nv2, w.mu_points,
                           w.kmeans_iters),
            "precision": "fp64",
            "kmeans_iters_run": r0["kmeans_iters"],
            "rk": rk,
            "rk_kmeans": r0["rk_kmeans"],
        },
        "zct_tflops": zct_tf,
        "zct_alone": {"tflops": zct_alone_tf, "frac_fp64_peak": zct_alone_tf / FP64_PEAK_TFLOPS,
                      "ms": avg_alone["ms_zct"],
                      "note": "FP64 ZC^T with no CUs reserved for the concurrent Cholesky "
                              "branch (ISDF_ZCT_RESERVE_CUS=0); zct_tflops is the in-solve rate"},
        "stage_ms": {k[3:]: round(v, 4) for k, v in avg.items()},
        "roofline": {
            "kernel": "k_potrs_dpp (blocked in-place triangular solves on the FP64 vector FMA "
                      "pipe, B operand broadcast by DPP; compute-bound: the published FP64 "
                      "vector and matrix peaks are both 78.6 TFLOP/s)",
            "bound": "mfma",
            "achieved": potrs_tf,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": potrs_tf / FP64_PEAK_TFLOPS,
            "traffic": traffic["bytes"] if traffic else None,
            "traffic_source": traffic["source"] if traffic else None,
            "algorithmic_flops_per_launch": fl["potrs"],
            "avg_launch_ms": avg["ms_potrs"],
        },
    }

    # Mixed forms on the same input, against the FP64 result of the timed run.
    if a.mixed:
        # Timed mixed solves first; the full-size fit errors (seconds of sustained FP64 MFMA
        # work at 96^3) come after every timed run, so they cannot heat or throttle it.
        theta_norm = torch.linalg.norm(d_theta)
        steps_m = max(1, a.steps // 2)
        runs = []
        for name, prec in (("bf16x3", isdf.ISDF_PREC_BF16X3), ("fp32", isdf.ISDF_PREC_FP32),
                           ("fp32+refine", isdf.ISDF_PREC_FP32 | isdf.ISDF_PREC_REFINE)):
            d_mix = torch.empty_like(d_theta)
            piv_m = np.zeros(w.ng, np.int32)
            dtm, avgm, _ = timed(prec, d_mix, piv_m, steps_m, 1)
            err = float(torch.linalg.norm(d_mix - d_theta) / theta_norm)
            runs.append((name, dtm, avgm, err, d_mix, piv_m))
        fit64, fit_ms = device_fit_error(torch, d_psi, d_theta, piv, dom, w)
        out["fit_error_fp64"] = fit64
        out["fit_error_kernel"] = {
            "what": "isdf_hip_fit_error over all Ng rows and K1*K2 pairs (FP64 MFMA residual GEMM)",
            "ms": fit_ms, "tflops": fl["fit"] / (fit_ms * 1e-3) / 1e12}
        mixed = []
        for name, dtm, avgm, err, d_mix, piv_m in runs:
            fitm, _ = device_fit_error(torch, d_psi, d_mix, piv_m, dom, w)
            mixed.append({
                "precision": name,
                "solves_per_s": steps_m / dtm,
                "speedup_vs_fp64": (steps_m / dtm) / (a.steps / dt),
                "theta_rel_err_vs_fp64": err,
                "fit_error": fitm,
                "fit_error_ratio_vs_fp64": fitm / fit64,
                "piv_equal_fp64": bool(np.array_equal(piv_m, piv)),
                "zct_tflops": fl["zct"] / (avgm["ms_zct"] * 1e-3) / 1e12,
                "theta_tflops": fl["potrs"] / (avgm["ms_potrs"] * 1e-3) / 1e12,
                # Fraction of the MFMA peak the Theta product runs on: FP32 counts its
                # nominal flops; BF16X3 runs six bf16 products per nominal multiply-add.
                "theta_frac_mfma_peak": (
                    fl["potrs"] / (avgm["ms_potrs"] * 1e-3) / 1e12 / FP32_PEAK_TFLOPS
                    if name.startswith("fp32") else
                    6.0 * fl["potrs"] / (avgm["ms_potrs"] * 1e-3) / 1e12 / BF16_PEAK_TFLOPS),
                "stage_ms": {k[3:]: round(v, 4) for k, v in avgm.items()},
            })
        del runs
        for m in mixed:
            m["meets_mixed_target"] = meets_target(m, fit64)
        out["mixed"] = mixed
        out["mixed_target"] = {
            "rule": ">= 3x the FP64 solve, Theta within 1e-6 relative of FP64, fit error "
                    "within 1.1x of FP64 (stated for the 96^3 workload)",
            "forms_meeting_it": [m["precision"] for m in mixed if m["meets_mixed_target"]],
            "fp64_zct_frac_of_peak": {"in_solve": zct_tf / FP64_PEAK_TFLOPS,
                                      "alone": zct_alone_tf / FP64_PEAK_TFLOPS,
                                      "target": 0.60}}
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(psi, w, r0["kmeans_iters"])
    print(json.dumps(out))


def bench_check(a):
    """check=true (checkcode.cpp:54-313) on the FP64 result of one solve:
    isdf_hip_check_device (fit + FFT Coulomb errors) timed over --steps calls,
    then the host validator once, on the same Theta and piv."""
    import torch
    import libisdf_amd as isdf
    from libisdf_amd import synth
    w = synth.WORKLOADS[a.workload]
    psi = synth.wavefunctions(w)
    rk = isdf.isdf_rank(w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points)
    d_psi = torch.from_numpy(psi).cuda()
    d_theta = torch.empty((rk, w.ng), dtype=torch.float64, device="cuda")
    piv = np.zeros(w.ng, np.int32)
    dom = isdf.Domain(w.length, w.ngrid)
    args = (w.nocc, w.nstate, w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points)
    isdf.solve(d_psi, dom, *args[:6], w.mu_points, d_theta, piv, w.kmeans_iters)
    for _ in range(a.warmup):
        isdf.check_device(d_psi, d_theta, piv, dom, *args)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        dev = isdf.check_device(d_psi, d_theta, piv, dom, *args)
    dt = (time.perf_counter() - t0) / a.steps
    th = d_theta.cpu().numpy()
    t0 = time.perf_counter()
    host = isdf.check(psi, th, piv, dom, *args)
    dt_host = time.perf_counter() - t0
    print(json.dumps({
        "what": "check=true on the device (isdf_hip_check_device) vs the host validator",
        "workload": w.name, "ng": w.ng, "pairs": (w.nv1 + w.nc1) * (w.nv2 + w.nc2), "rk": rk,
        "device_ms": dt * 1e3, "host_s": dt_host, "device": dev, "host": host,
        "coulomb_rel_diff": abs(dev["coulomb_error"] - host["coulomb_error"])
        / host["coulomb_error"],
        "fit_rel_diff": abs(dev["fit_error"] - host["fit_error"]) / host["fit_error"]}))


def bench_batch(a):
    """The uneven batch: 24 instances (grid n in {24..64}, nstate in 16..96) solved
    by one batched call (FP32 MFMA contractions, FP64 Cholesky chain, grouped
    launches replayed from a hipGraph) against a loop of single FP64 solves."""
    import torch
    import libisdf_amd as isdf
    from libisdf_amd import synth

    ws = synth.batch5()
    dev = torch.device("cuda:0")
    inst, ref = [], []
    for w in ws:
        psi = synth.wavefunctions(w)
        rk = isdf.isdf_rank(w.nv1, w.nc1, w.nv2, w.nc2, w.mu_points)
        d_psi = torch.from_numpy(psi).to(dev)
        inst.append(dict(d_psi=d_psi, domain=isdf.Domain(w.length, w.ngrid), nocc=w.nocc,
                         nstate=w.nstate, nv1=w.nv1, nc1=w.nc1, nv2=w.nv2, nc2=w.nc2,
                         mu_points=w.mu_points, max_iter=w.kmeans_iters,
                         d_theta=torch.empty((rk, w.ng), dtype=torch.float64, device=dev),
                         piv=np.zeros(w.ng, np.int32)))
        ref.append(torch.empty((rk, w.ng), dtype=torch.float64, device=dev))
    stream = torch.cuda.current_stream()

    def loop_fp64():
        for x, r in zip(inst, ref):
            isdf.solve(x["d_psi"], x["domain"], x["nocc"], x["nstate"], x["nv1"], x["nc1"],
                       x["nv2"], x["nc2"], x["mu_points"], r, x["piv"], x["max_iter"],
                       stream=stream)

    for _ in range(max(1, a.warmup)):
        loop_fp64()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        loop_fp64()
    torch.cuda.synchronize()
    dt64 = (time.perf_counter() - t0) / a.steps
    piv64 = [x["piv"].copy() for x in inst]

    out_modes = {}
    for name, prec in (("fp32", isdf.ISDF_PREC_FP32), ("bf16x3", isdf.ISDF_PREC_BF16X3)):
        _, first = isdf.solve_batched(inst, precision=prec, stream=stream)  # tune + capture
        for _ in range(max(0, a.warmup - 1)):
            isdf.solve_batched(inst, precision=prec, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        brs = []
        for _ in range(a.steps):
            brs.append(isdf.solve_batched(inst, precision=prec, stream=stream)[1])
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / a.steps
        errs = [float(torch.linalg.norm(x["d_theta"] - r) / torch.linalg.norm(r))
                for x, r in zip(inst, ref)]
        pv = all(np.array_equal(x["piv"], p) for x, p in zip(inst, piv64))
        theta_flops = sum(2.0 * w.ng * w.mu_points ** 2 for w in ws)
        tv = first["theta_tile"]
        out_modes[name] = dict(
            solves_per_s=len(ws) / dt, ms_per_batch=dt * 1e3,
            speedup_vs_fp64_loop=dt64 / dt, max_theta_rel_err_vs_fp64=max(errs),
            piv_equal_fp64=bool(pv), graph_replayed=bool(brs[-1]["graph_replayed"]),
            tiles=dict(zct=first["zct_tile"], theta=tv, zct_ms=first["zct_tile_ms"][:3],
                       theta_ms=first["theta_tile_ms"][:3]),
            theta_grouped_tflops=theta_flops / (first["theta_tile_ms"][tv] * 1e-3) / 1e12)
    best = out_modes["fp32"]
    out = {
        "metric": METRIC,
        "value": best["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": best["ms_per_batch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded cosine plane waves, FP64 CholeskyQR2-orthonormalised)",
        "config": {"workload": "cfg5: 24 uneven instances (grid n in {24,32,40,48,64}, "
                               "nstate in 16..96, rk = 4 nstate, 30 K-means iterations), one "
                               "batched call per step, FP32 MFMA contractions, FP64 Cholesky",
                   "instances": len(ws), "total_ng": sum(w.ng for w in ws)},
        "fp64_loop_solves_per_s": len(ws) / dt64,
        "batched": out_modes,
        "roofline": {
            "kernel": "k_theta_lp_grouped (persistent grouped Theta GEMM, FP32 MFMA)",
            "bound": "mfma", "achieved": best["theta_grouped_tflops"], "peak": 157.3,
            "unit": "TFLOP/s", "frac": best["theta_grouped_tflops"] / 157.3, "traffic": None},
    }
    print(json.dumps(out))


def device_fit_error(torch, d_psi, d_theta, piv, dom, w):
    """||Z - Theta C||_F / ||Z||_F over every grid row and pair
    (src/common/checkcode.cpp:62-166), on the device; returns (fit, ms)."""
    import libisdf_amd as isdf
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = isdf.fit_error(d_psi, d_theta, piv, dom, w.nocc, w.nstate, w.nv1, w.nc1, w.nv2, w.nc2,
                       w.mu_points)
    return r["fit_error"], (time.perf_counter() - t0) * 1e3


if __name__ == "__main__":
    main()
