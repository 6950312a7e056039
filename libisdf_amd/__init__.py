"""libisdf_amd — MI355X-native ISDF hot path of LibISDF.

Python mirror of the reference's GPU interface (include/isdf.hpp:37):

    ISDF(d_psirow, domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points,
         d_thetaCol, piv, KmeansMaxIter_ISDF, check=False)

The work happens in libisdf_hip.so (hand-written HIP kernels for gfx950 behind
the C ABI of include/isdf_hip.h); this module only passes device pointers and
sizes.  PyTorch supplies device memory and the current stream.  There is no
CPU fallback: without the shared library every call raises.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (ISDF_EHIP, ISDF_EINVAL, ISDF_ENOMEM, ISDF_ENOTSPD, ISDF_OK,
                   ISDF_PREC_BF16X3, ISDF_PREC_FP32, ISDF_PREC_FP64, ISDF_PREC_REFINE,
                   ISDF_SELECT_KMEANS, ISDF_SELECT_QRCP)

__all__ = ["Domain", "ISDF", "ISDF_host", "isdf_rank", "buffer_sizes", "solve", "solve_noalloc",
           "solve_host",
           "solve_batched",
           "check", "check_device", "fit_error", "ISDF_SELECT_KMEANS", "ISDF_SELECT_QRCP", "ISDF_PREC_FP64", "ISDF_PREC_FP32",
           "ISDF_PREC_BF16X3", "ISDF_PREC_REFINE"]


class Domain:
    """Simulation cell, as include/common/domain.hpp:8-36 (one rank: no communicators)."""

    def __init__(self, length=(0.0, 0.0, 0.0), numGrid=(0, 0, 0), posStart=(0.0, 0.0, 0.0),
                 numGridFine=(0, 0, 0)):
        self.length = tuple(float(x) for x in length)
        self.posStart = tuple(float(x) for x in posStart)
        self.numGrid = tuple(int(x) for x in numGrid)
        self.numGridFine = tuple(int(x) for x in numGridFine)

    def Volume(self):
        return self.length[0] * self.length[1] * self.length[2]

    def NumGridTotal(self):
        return self.numGrid[0] * self.numGrid[1] * self.numGrid[2]


def _problem(domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, max_iter,
             precision=ISDF_PREC_FP64, ld_psi=0, select=ISDF_SELECT_KMEANS):
    p = _lib.IsdfProblem()
    for d in range(3):
        p.ngrid[d] = int(domain.numGrid[d])
        p.length[d] = float(domain.length[d])
    p.nocc, p.nstate = int(nocc), int(nstate)
    p.nv1, p.nc1, p.nv2, p.nc2 = int(nv1), int(nc1), int(nv2), int(nc2)
    p.mu_points = int(mu_points)
    p.kmeans_max_iter = int(max_iter)
    p.precision = int(precision)
    p.select = int(select)
    p.ld_psi = int(ld_psi)
    return p


def isdf_rank(nv1, nc1, nv2, nc2, mu_points):
    """rk of the reference's rank rule (src/common/isdf.cpp:160-164)."""
    p = _lib.IsdfProblem()
    p.nv1, p.nc1, p.nv2, p.nc2, p.mu_points = nv1, nc1, nv2, nc2, mu_points
    return _lib.lib().isdf_hip_rank(ctypes.byref(p))


def buffer_sizes(ng, nstate, rk, ld_psi=0):
    """Minimum element counts of the device buffers of one solve: psi is the
    column-major Ng x nstate matrix with leading dimension ld (0 -> Ng), so it
    spans ld*(nstate-1) + Ng doubles; Theta is Ng x rk with ld Ng."""
    ld = ld_psi if ld_psi else ng
    if ld < ng:
        raise ValueError("ld_psi (%d) is below Ng (%d)" % (ld, ng))
    return ld * (max(nstate, 1) - 1) + ng, ng * rk


def _dev_ptr(x, need=None, name="buffer", current=None):
    """Device address of x.  For a tensor: on the current HIP device, float64,
    contiguous and holding at least `need` elements (a short buffer would be
    written or read past its end by the kernels).  An int is taken as a raw
    device address whose size the caller vouches for.  `current`: the current
    device index when the caller already has it."""
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        import torch
        if not x.is_cuda:
            raise ValueError("%s: device buffer expected (a CUDA/HIP tensor)" % name)
        if x.dtype != torch.float64:
            raise ValueError("%s: float64 buffer expected" % name)
        if not x.is_contiguous():
            raise ValueError("%s: contiguous buffer expected" % name)
        cur = torch.cuda.current_device() if current is None else current
        if x.device.index != cur:
            raise ValueError("%s is on device %d, the current device is %d"
                             % (name, x.device.index, cur))
        if need is not None and x.numel() < need:
            raise ValueError("%s holds %d elements, the problem needs %d"
                             % (name, x.numel(), need))
        return x.data_ptr()
    raise TypeError("device buffer must be a torch tensor on the GPU or an int pointer")


def _current_device(instances):
    """The current HIP device index, read once for a batch (None when no
    instance holds a tensor)."""
    if any(hasattr(a.get("d_psi"), "data_ptr") or hasattr(a.get("d_theta"), "data_ptr")
           for a in instances):
        import torch
        return torch.cuda.current_device()
    return None


def _host_piv(piv, n, name="piv"):
    if not (isinstance(piv, np.ndarray) and piv.dtype == np.int32 and piv.flags.c_contiguous
            and piv.size >= n):
        raise ValueError("%s must be a contiguous numpy int32 array of at least %d entries"
                         % (name, n))
    return piv


def _stream_ptr(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def _raise(rc):
    msg = _lib.last_error()
    if rc == ISDF_EINVAL:
        raise ValueError(msg)
    if rc == ISDF_ENOMEM:
        raise MemoryError(msg)
    # ISDF_ENOTSPD mirrors ErrorHandling("Matrix is not positive definite."),
    # which throws std::runtime_error (src/common/global.cpp:7-11).
    raise RuntimeError(msg)


def solve(d_psi, domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, d_theta, piv,
          max_iter, stream=None, precision=ISDF_PREC_FP64, debug=None, ld_psi=0,
          select=ISDF_SELECT_KMEANS):
    """Run the C-ABI solve.  piv: numpy int32 array of Ng entries (written).
    debug: optional dict with any of 'trace' (numpy int32 array), 'weights',
    'labels', 'cct' (device tensors).  select: ISDF_SELECT_KMEANS (the GPU
    ISDF's method) or ISDF_SELECT_QRCP.  Returns the report as a dict."""
    L = _lib.lib()
    p = _problem(domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, max_iter, precision, ld_psi,
                 select)
    ng = domain.NumGridTotal()
    _host_piv(piv, ng)
    rk = isdf_rank(nv1, nc1, nv2, nc2, mu_points)
    n_psi, n_theta = buffer_sizes(ng, nstate, rk, ld_psi)
    psi_ptr = _dev_ptr(d_psi, n_psi, "d_psi")
    theta_ptr = _dev_ptr(d_theta, n_theta, "d_theta")
    rep = _lib.IsdfReport()
    dbg = None
    if debug:
        dbg = _lib.IsdfDebug()
        if "trace" in debug:
            t = debug["trace"]
            assert t.dtype == np.int32 and t.size >= max(max_iter, 1)
            dbg.h_trace = t.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        import torch
        for key, dt, n in (("weights", torch.float64, ng), ("labels", torch.int32, ng),
                           ("cct", torch.float64, rk * rk)):
            if key in debug:
                x = debug[key]
                if not (x.is_cuda and x.dtype == dt and x.is_contiguous() and x.numel() >= n
                        and x.device.index == torch.cuda.current_device()):
                    raise ValueError("debug['%s'] must be a contiguous %s tensor of at least %d "
                                     "elements on the current device" % (key, dt, n))
                setattr(dbg, "d_" + key, x.data_ptr())
    rc = L.isdf_hip_solve(ctypes.byref(p), ctypes.c_void_p(psi_ptr), ctypes.c_void_p(theta_ptr),
                          piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                          ctypes.c_void_p(_stream_ptr(stream)), ctypes.byref(rep),
                          ctypes.byref(dbg) if dbg is not None else None)
    if rc != ISDF_OK:
        _raise(rc)
    return rep.as_dict()


def solve_noalloc(d_psi, domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, d_theta, piv,
                  max_iter, stream=None, precision=ISDF_PREC_FP64, ld_psi=0):
    """isdf_hip_solve_noalloc: the solve of solve() with no device or host
    allocation (workspace in a static 256 MiB device array, every stage on
    `stream`, K-means selection only, no debug outputs).  Theta and piv equal
    solve()'s bit for bit.  Returns the report as a dict (stage times 0)."""
    L = _lib.lib()
    p = _problem(domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, max_iter, precision, ld_psi)
    ng = domain.NumGridTotal()
    _host_piv(piv, ng)
    rk = isdf_rank(nv1, nc1, nv2, nc2, mu_points)
    n_psi, n_theta = buffer_sizes(ng, nstate, rk, ld_psi)
    psi_ptr = _dev_ptr(d_psi, n_psi, "d_psi")
    theta_ptr = _dev_ptr(d_theta, n_theta, "d_theta")
    rep = _lib.IsdfReport()
    rc = L.isdf_hip_solve_noalloc(ctypes.byref(p), ctypes.c_void_p(psi_ptr),
                                  ctypes.c_void_p(theta_ptr),
                                  piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                  ctypes.c_void_p(_stream_ptr(stream)), ctypes.byref(rep))
    if rc != ISDF_OK:
        _raise(rc)
    return rep.as_dict()


def solve_host(psi, domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, theta, piv, max_iter,
               precision=ISDF_PREC_FP64, debug=None, stream=0, select=ISDF_SELECT_KMEANS):
    """isdf_hip_solve_host: the solve of solve() on host buffers.  psi: numpy
    float64 (nstate, Ng) C-contiguous (the column-major Ng x nstate matrix);
    theta: numpy float64 (rk, Ng) C-contiguous, written; piv as in solve().
    debug: optional dict with 'trace' (numpy int32 array).  stream: a
    hipStream_t as an int, 0 for the default stream."""
    L = _lib.lib()
    p = _problem(domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, max_iter, precision,
                 select=select)
    ng = domain.NumGridTotal()
    rk = isdf_rank(nv1, nc1, nv2, nc2, mu_points)
    for name, a, n in (("psi", psi, ng * max(nstate, 0)), ("theta", theta, ng * max(rk, 0))):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
                and a.size >= n):
            raise ValueError("%s must be a contiguous numpy float64 array of %d entries"
                             % (name, n))
    if not (isinstance(piv, np.ndarray) and piv.dtype == np.int32 and piv.flags.c_contiguous
            and piv.size >= ng):
        raise ValueError("piv must be a contiguous numpy int32 array of Ng entries")
    if theta.flags.writeable is False or piv.flags.writeable is False:
        raise ValueError("theta and piv must be writeable")
    rep = _lib.IsdfReport()
    dbg = None
    if debug and "trace" in debug:
        dbg = _lib.IsdfDebug()
        t = debug["trace"]
        assert t.dtype == np.int32 and t.size >= max(max_iter, 1)
        dbg.h_trace = t.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.isdf_hip_solve_host(ctypes.byref(p), psi.ctypes.data_as(dp), theta.ctypes.data_as(dp),
                               piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                               ctypes.c_void_p(stream), ctypes.byref(rep),
                               ctypes.byref(dbg) if dbg is not None else None)
    if rc != ISDF_OK:
        _raise(rc)
    return rep.as_dict()


def solve_batched(instances, precision=ISDF_PREC_FP32, stream=None):
    """Many independent instances in one call (isdf_hip_solve_batched; no
    counterpart in the reference).  instances: list of dicts with keys d_psi,
    domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, d_theta, piv,
    max_iter (as in solve()).  Returns (list of per-instance reports, batch
    report)."""
    L = _lib.lib()
    n = len(instances)
    cur = _current_device(instances)
    probs = (_lib.IsdfProblem * n)()
    psis = (ctypes.c_void_p * n)()
    thetas = (ctypes.c_void_p * n)()
    pivs = (ctypes.POINTER(ctypes.c_int) * n)()
    for i, a in enumerate(instances):
        probs[i] = _problem(a["domain"], a["nocc"], a["nstate"], a["nv1"], a["nc1"], a["nv2"],
                            a["nc2"], a["mu_points"], a["max_iter"], precision)
        ng = a["domain"].NumGridTotal()
        piv = _host_piv(a["piv"], ng, "instance %d piv" % i)
        rk = isdf_rank(a["nv1"], a["nc1"], a["nv2"], a["nc2"], a["mu_points"])
        n_psi, n_theta = buffer_sizes(ng, a["nstate"], rk)
        psis[i] = _dev_ptr(a["d_psi"], n_psi, "instance %d d_psi" % i, cur)
        thetas[i] = _dev_ptr(a["d_theta"], n_theta, "instance %d d_theta" % i, cur)
        pivs[i] = piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    reps = (_lib.IsdfReport * n)()
    brep = _lib.IsdfBatchReport()
    rc = L.isdf_hip_solve_batched(n, probs, psis, thetas, pivs,
                                  ctypes.c_void_p(_stream_ptr(stream)), reps, ctypes.byref(brep))
    if rc != ISDF_OK:
        _raise(rc)
    return [r.as_dict() for r in reps], brep.as_dict()


def check(psi, theta, piv, domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, coulomb=True):
    """check=true validator (src/common/checkcode.cpp:54-313) on host arrays:
    psi (nstate, Ng), theta (rk, Ng) float64, piv int32.  Returns dict."""
    L = _lib.lib()
    p = _problem(domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, 0)
    psi = np.ascontiguousarray(psi, np.float64)
    theta = np.ascontiguousarray(theta, np.float64)
    piv = np.ascontiguousarray(piv, np.int32)
    out = _lib.IsdfCheck()
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.isdf_hip_check(ctypes.byref(p), psi.ctypes.data_as(dp), theta.ctypes.data_as(dp),
                          piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(bool(coulomb)),
                          ctypes.byref(out))
    if rc != ISDF_OK:
        _raise(rc)
    return dict(fit_error=out.fit_error, z_norm=out.z_norm, coulomb_error=out.coulomb_error,
                coulomb_norm=out.coulomb_norm)


_validate = check


def _check_args(d_psi, d_theta, piv, domain, nstate, nv1, nc1, nv2, nc2, mu_points, ld_psi):
    """Device addresses and the piv array of a check / fit-error call, with
    the buffer sizes of buffer_sizes() and at least rk pivots."""
    ng = domain.NumGridTotal()
    rk = isdf_rank(nv1, nc1, nv2, nc2, mu_points)
    n_psi, n_theta = buffer_sizes(ng, nstate, rk, ld_psi)
    piv = np.ascontiguousarray(piv, np.int32)
    if piv.size < rk:
        raise ValueError("piv holds %d entries, rk = %d are needed" % (piv.size, rk))
    return _dev_ptr(d_psi, n_psi, "d_psi"), _dev_ptr(d_theta, n_theta, "d_theta"), piv


def check_device(d_psi, d_theta, piv, domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points,
                 coulomb=True, stream=None, ld_psi=0):
    """isdf_hip_check_device: check=true on the device (checkcode.cpp:54-313),
    the fit error and, with coulomb=True, the FFT Coulomb error.  Arguments as
    fit_error().  Returns the dict check() returns."""
    L = _lib.lib()
    p = _problem(domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, 0, ISDF_PREC_FP64, ld_psi)
    psi_ptr, theta_ptr, piv = _check_args(d_psi, d_theta, piv, domain, nstate, nv1, nc1, nv2,
                                          nc2, mu_points, ld_psi)
    out = _lib.IsdfCheck()
    rc = L.isdf_hip_check_device(ctypes.byref(p), ctypes.c_void_p(psi_ptr),
                                 ctypes.c_void_p(theta_ptr),
                                 piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                 int(bool(coulomb)), ctypes.c_void_p(_stream_ptr(stream)),
                                 ctypes.byref(out))
    if rc != ISDF_OK:
        _raise(rc)
    return dict(fit_error=out.fit_error, z_norm=out.z_norm, coulomb_error=out.coulomb_error,
                coulomb_norm=out.coulomb_norm)


def fit_error(d_psi, d_theta, piv, domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points,
              stream=None, ld_psi=0):
    """isdf_hip_fit_error: ||Z - Theta C||_F / ||Z||_F on the device over all
    rows and pairs (the fit-error half of check=true, checkcode.cpp:62-166).
    d_psi, d_theta: device float64 tensors as solve() takes and writes them;
    piv: host int32 numpy array (first rk entries used).  Returns dict with
    fit_error and z_norm."""
    L = _lib.lib()
    p = _problem(domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, 0, ISDF_PREC_FP64, ld_psi)
    psi_ptr, theta_ptr, piv = _check_args(d_psi, d_theta, piv, domain, nstate, nv1, nc1, nv2,
                                          nc2, mu_points, ld_psi)
    out = _lib.IsdfCheck()
    rc = L.isdf_hip_fit_error(ctypes.byref(p), ctypes.c_void_p(psi_ptr),
                              ctypes.c_void_p(theta_ptr),
                              piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              ctypes.c_void_p(_stream_ptr(stream)), ctypes.byref(out))
    if rc != ISDF_OK:
        _raise(rc)
    return dict(fit_error=out.fit_error, z_norm=out.z_norm)


def ISDF_host(psirow, domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, thetaCol, piv,
              KmeansMaxIter_ISDF, scalblocksize=64, s="scalapack", s1="Kmeans", check=False):
    """Mirror of the CPU-interface isdf::isdffunc::ISDF (include/isdf.hpp:39,
    src/common/isdf.cpp:8-127) at one rank, backed by the GPU solve.

    psirow: host float64 numpy array, the column-major Ng x nstate matrix
    ((nstate, Ng) C-contiguous).  thetaCol: host float64 numpy array (rk, Ng),
    written.  piv: host int32 numpy array of Ng entries, zeroed and written.
    s: "scalapack" or "MPI" (both give the same Theta at one rank; the
    reference's two basis routes differ only in their distribution).  Any other
    s does nothing after zeroing piv, as in the reference.  s1: "Kmeans" or
    "QRCP" (ISDF_SELECT_QRCP); anything else raises "Wrong select points
    method" (isdf.cpp:73, :112).  scalblocksize is the ScaLAPACK block size of the
    reference; it has no role in the one-rank solve and must be positive.
    """
    ng = domain.NumGridTotal()
    if not (isinstance(piv, np.ndarray) and piv.dtype == np.int32 and piv.size >= ng):
        raise ValueError("piv must be a numpy int32 array of Ng entries")
    piv[:ng] = 0
    if s not in ("scalapack", "MPI"):
        return {}
    if s1 not in ("Kmeans", "QRCP"):
        raise RuntimeError("Wrong select points method")
    if scalblocksize <= 0:
        raise ValueError("scalblocksize must be positive")
    sel = ISDF_SELECT_QRCP if s1 == "QRCP" else ISDF_SELECT_KMEANS
    rep = solve_host(psirow, domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, thetaCol, piv,
                     KmeansMaxIter_ISDF, select=sel)
    if check:
        rep.update(_validate(psirow, thetaCol, piv, domain, nocc, nstate, nv1, nc1, nv2, nc2,
                             mu_points))
    return rep


def ISDF(d_psirow, domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, d_thetaCol, piv,
         KmeansMaxIter_ISDF, check=False, stream=None):
    """Drop-in mirror of isdf::isdffunc::ISDF (GPU, include/isdf.hpp:37).
    (ISDF_host mirrors the CPU interface of include/isdf.hpp:39.)

    d_psirow: device float64 tensor holding the column-major Ng x nstate matrix
    (a (nstate, Ng) C-contiguous tensor).  d_thetaCol: device float64 tensor of
    Ng x rk column-major ((rk, Ng) C-contiguous), written.  piv: host int32
    numpy array of Ng entries, written.  Raises RuntimeError with the
    reference's message when the C C^T Cholesky fails.  With check=True the
    fit and Coulomb errors are computed on the device (check_device) and
    returned in the report (keys 'fit_error', 'coulomb_error'), as checkcode
    prints them.
    """
    rep = solve(d_psirow, domain, nocc, nstate, nv1, nc1, nv2, nc2, mu_points, d_thetaCol, piv,
                KmeansMaxIter_ISDF, stream=stream)
    if check:
        # Both halves of checkcode on the device, on the caller's buffers (as
        # the C++ drop-in does): no host copy of psi or Theta.
        rep.update(check_device(d_psirow, d_thetaCol, piv, domain, nocc, nstate, nv1, nc1, nv2,
                                nc2, mu_points, coulomb=True, stream=stream))
    return rep
