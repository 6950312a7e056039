"""Build libisdf_hip.so (HIP kernels + C ABI + C++ drop-in) for gfx950 in-tree.

    python -m libisdf_amd.build            # compile everything
    python -m libisdf_amd.build --verbose  # also print kernel resource usage
    python -m libisdf_amd.build --sanitize # host code with ASan + UBSan into _san/

The shared library lands next to this file (libisdf_amd/libisdf_hip.so) so it
travels with the source tree.  hipcc cross-compiles; no GPU is needed.
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libisdf_hip.so")
# Host-sanitizer build: the host half of every file with AddressSanitizer and
# UndefinedBehaviorSanitizer (-Xarch_host: the device code is built as usual).
# For CPU-only test runs with the clang ASan runtime preloaded.
SAN_BUILD = os.path.join(HERE, "_build_san")
SAN_LIB = os.path.join(HERE, "_san", "libisdf_hip.so")
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g"]
# Register-only arithmetic probes (bench.py --peak-probe); not part of the ISDF library.
PROBE_LIB = os.path.join(HERE, "libisdf_probe.so")

ARCH = os.environ.get("ISDF_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I" + os.path.join(ROOT, "include"),
]

# Per-source extra flags.  The K-means file must not contract a*b+c into FMA:
# its results are compared bit for bit with the CPU reference.
SOURCES = {
    "kmeans.hip": ["-ffp-contract=off"],
    "basis.hip": [],
    "lowprec.hip": [],
    "fit.hip": [],
    "coulomb.hip": [],
    "qrcp.hip": [],
    "solve.cpp": ["-x", "hip"],
    "batch.cpp": ["-x", "hip"],
    "check.cpp": [],
    "dropin.cpp": ["-x", "hip"],
}


def _compile(src, extra, verbose, build_dir=BUILD):
    obj = os.path.join(build_dir, src + ".o")
    cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose and src.endswith(".hip"):
        cmd += ["-Rpass-analysis=kernel-resource-usage"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = r.stdout + r.stderr
    lines = [l for l in out.splitlines()
             if "packed-fp32-ops" not in l or "error" in l]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, "\n".join(lines)))
    return obj, "\n".join(lines)


def _link(objs, lib, extra=()):
    tmp = lib + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=" + ARCH, "-o", tmp] + list(extra) + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, lib)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))] + ["probe.hip"]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        res = list(ex.map(lambda s: _compile(s, SOURCES.get(s, []), verbose), srcs))
    objs = []
    for (obj, log), s in zip(res, srcs):
        if s != "probe.hip":
            objs.append(obj)
        if verbose and log.strip():
            print("== %s\n%s" % (s, log))
    _link(objs, LIB)
    _link([os.path.join(BUILD, "probe.hip.o")], PROBE_LIB)
    return LIB


def build_sanitized():
    """The library with the host code under ASan + UBSan (SAN_LIB)."""
    os.makedirs(SAN_BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(SAN_LIB), exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        res = list(ex.map(lambda s: _compile(s, SOURCES[s] + SAN_FLAGS, False, SAN_BUILD), srcs))
    _link([obj for obj, _ in res], SAN_LIB, ["-Xarch_host", "-shared-libsan",
                                             "-Xarch_host", "-fsanitize=address,undefined"])
    return SAN_LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--sanitize", action="store_true")
    a = ap.parse_args()
    print(build_sanitized() if a.sanitize else build(a.verbose))
    sys.exit(0)
