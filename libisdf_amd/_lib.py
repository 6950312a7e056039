"""ctypes binding of the C ABI in include/isdf_hip.h (libisdf_amd/libisdf_hip.so).

There is no fallback: if the shared library is missing or does not load, every
entry point raises.  Build it with ``python -m libisdf_amd.build``.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# ISDF_HIP_LIB selects another build of the same library (the host-sanitizer
# build, libisdf_amd/_san/libisdf_hip.so, for CPU-only test runs).
LIB_PATH = os.environ.get("ISDF_HIP_LIB") or os.path.join(_HERE, "libisdf_hip.so")

ISDF_OK, ISDF_EINVAL, ISDF_ENOTSPD, ISDF_EHIP, ISDF_ENOMEM = 0, 1, 2, 3, 4
ISDF_PREC_FP64, ISDF_PREC_FP32, ISDF_PREC_BF16X3, ISDF_PREC_REFINE = 0, 1, 2, 16
ISDF_SELECT_KMEANS, ISDF_SELECT_QRCP = 0, 1

# Every symbol the public header declares; tests check the library exports them.
EXPORTS = (
    "isdf_hip_solve",
    "isdf_hip_solve_host",
    "isdf_hip_solve_noalloc",
    "isdf_hip_rank",
    "isdf_hip_last_error",
    "isdf_hip_version",
    "isdf_hip_release_workspace",
    "isdf_hip_check",
    "isdf_hip_fit_error",
    "isdf_hip_check_device",
    "isdf_hip_solve_batched",
)


class IsdfProblem(ctypes.Structure):
    _fields_ = [
        ("ngrid", ctypes.c_int * 3),
        ("length", ctypes.c_double * 3),
        ("nocc", ctypes.c_int),
        ("nstate", ctypes.c_int),
        ("nv1", ctypes.c_int),
        ("nc1", ctypes.c_int),
        ("nv2", ctypes.c_int),
        ("nc2", ctypes.c_int),
        ("mu_points", ctypes.c_int),
        ("kmeans_max_iter", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("select", ctypes.c_int),
        ("ld_psi", ctypes.c_longlong),
    ]


class IsdfReport(ctypes.Structure):
    _fields_ = [
        ("rk", ctypes.c_int),
        ("rk_kmeans", ctypes.c_int),
        ("kmeans_iters", ctypes.c_int),
        ("npt", ctypes.c_int),
        ("potrf_info", ctypes.c_int),
        ("reserved", ctypes.c_int),
        ("eps", ctypes.c_double),
        ("ms_weights", ctypes.c_double),
        ("ms_kmeans", ctypes.c_double),
        ("ms_select", ctypes.c_double),
        ("ms_gather", ctypes.c_double),
        ("ms_zct", ctypes.c_double),
        ("ms_cct", ctypes.c_double),
        ("ms_potrf", ctypes.c_double),
        ("ms_potrs", ctypes.c_double),
        ("ms_total", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class IsdfDebug(ctypes.Structure):
    _fields_ = [
        ("h_trace", ctypes.POINTER(ctypes.c_int)),
        ("d_weights", ctypes.c_void_p),
        ("d_labels", ctypes.c_void_p),
        ("d_cct", ctypes.c_void_p),
    ]


class IsdfBatchReport(ctypes.Structure):
    _fields_ = [
        ("ms_total", ctypes.c_double),
        ("graph_replayed", ctypes.c_int),
        ("zct_tile", ctypes.c_int),
        ("theta_tile", ctypes.c_int),
        ("reserved", ctypes.c_int),
        ("zct_tile_ms", ctypes.c_double * 4),
        ("theta_tile_ms", ctypes.c_double * 4),
    ]

    def as_dict(self):
        return dict(ms_total=self.ms_total, graph_replayed=self.graph_replayed,
                    zct_tile=self.zct_tile, theta_tile=self.theta_tile,
                    zct_tile_ms=list(self.zct_tile_ms), theta_tile_ms=list(self.theta_tile_ms))


class IsdfCheck(ctypes.Structure):
    _fields_ = [
        ("fit_error", ctypes.c_double),
        ("z_norm", ctypes.c_double),
        ("coulomb_error", ctypes.c_double),
        ("coulomb_norm", ctypes.c_double),
    ]


_lib = None


def lib():
    """Load libisdf_hip.so (raises if it is absent: there is no CPU path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libisdf_hip.so is not built (%s); run `python -m libisdf_amd.build`"
                           % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.isdf_hip_solve.restype = ctypes.c_int
    L.isdf_hip_solve.argtypes = [ctypes.POINTER(IsdfProblem), ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.POINTER(ctypes.c_int), ctypes.c_void_p,
                                 ctypes.POINTER(IsdfReport), ctypes.POINTER(IsdfDebug)]
    L.isdf_hip_solve_noalloc.restype = ctypes.c_int
    L.isdf_hip_solve_noalloc.argtypes = [ctypes.POINTER(IsdfProblem), ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                         ctypes.c_void_p, ctypes.POINTER(IsdfReport)]
    L.isdf_hip_solve_host.restype = ctypes.c_int
    L.isdf_hip_solve_host.argtypes = [ctypes.POINTER(IsdfProblem), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_int), ctypes.c_void_p,
                                      ctypes.POINTER(IsdfReport), ctypes.POINTER(IsdfDebug)]
    L.isdf_hip_rank.restype = ctypes.c_int
    L.isdf_hip_rank.argtypes = [ctypes.POINTER(IsdfProblem)]
    L.isdf_hip_last_error.restype = ctypes.c_char_p
    L.isdf_hip_last_error.argtypes = []
    L.isdf_hip_version.restype = ctypes.c_char_p
    L.isdf_hip_version.argtypes = []
    L.isdf_hip_release_workspace.restype = ctypes.c_int
    L.isdf_hip_release_workspace.argtypes = []
    L.isdf_hip_check.restype = ctypes.c_int
    L.isdf_hip_check.argtypes = [ctypes.POINTER(IsdfProblem), ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                                 ctypes.c_int, ctypes.POINTER(IsdfCheck)]
    L.isdf_hip_fit_error.restype = ctypes.c_int
    L.isdf_hip_fit_error.argtypes = [ctypes.POINTER(IsdfProblem), ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                     ctypes.c_void_p, ctypes.POINTER(IsdfCheck)]
    L.isdf_hip_check_device.restype = ctypes.c_int
    L.isdf_hip_check_device.argtypes = [ctypes.POINTER(IsdfProblem), ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                        ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(IsdfCheck)]
    L.isdf_hip_solve_batched.restype = ctypes.c_int
    L.isdf_hip_solve_batched.argtypes = [
        ctypes.c_int, ctypes.POINTER(IsdfProblem), ctypes.POINTER(ctypes.c_void_p),
        ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.POINTER(ctypes.c_int)),
        ctypes.c_void_p, ctypes.POINTER(IsdfReport), ctypes.POINTER(IsdfBatchReport)]
    _lib = L
    return L


def last_error():
    return lib().isdf_hip_last_error().decode()
