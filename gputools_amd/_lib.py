"""ctypes binding of libgputools_amd.so (the C-ABI in include/gputools_amd.h).

The library is the product: there is no CPU fallback.  If it has not been
built, every entry point raises :class:`LibraryNotBuilt`.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgputools_amd.so")

# enum values of include/gputools_amd.h
EUCLIDEAN, MAXIMUM, MANHATTAN, CANBERRA, BINARY, MINKOWSKI = range(6)
METHOD_NAMES = ("euclidean", "maximum", "manhattan", "canberra", "binary", "minkowski")
PREC_FP32, PREC_BF16X3, PREC_FP16X3 = 0, 1, 2
PRECISIONS = {"fp32": PREC_FP32, "bf16x3": PREC_BF16X3, "fp16x3": PREC_FP16X3}
LAYOUT_PACKED, LAYOUT_FULL = 0, 1
# GT_NOALLOC_SCRATCH_BYTES: the static scratch of gt_dist_noalloc, per GPU
NOALLOC_SCRATCH_BYTES = 64 << 20
# gt_hc_method (hc_method order, src/hcluster.h:5-16); names as getClusterEnum spells them
HC_METHOD_NAMES = ("single", "complete", "wpgma", "average", "median", "centroid",
                   "flexible_group", "flexible", "ward", "mcquitty")
USE_EVERYTHING, USE_PAIRWISE = 0, 1
GT_OK, GT_ERR_INVALID, GT_ERR_HIP, GT_ERR_ALLOC, GT_ERR_UNSUPPORTED, GT_ERR_WORKSPACE = range(6)
STATUS_NAMES = {0: "GT_OK", 1: "GT_ERR_INVALID", 2: "GT_ERR_HIP", 3: "GT_ERR_ALLOC",
                4: "GT_ERR_UNSUPPORTED", 5: "GT_ERR_WORKSPACE"}


class LibraryNotBuilt(RuntimeError):
    pass


class GputoolsError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), message))
        self.status = status


_lock = threading.Lock()
_lib = None

c_i64 = ctypes.c_int64
c_p = ctypes.c_void_p
c_int = ctypes.c_int
c_float = ctypes.c_float
c_size = ctypes.c_size_t
c_fp = ctypes.POINTER(ctypes.c_float)
c_ip = ctypes.POINTER(ctypes.c_int)
c_cpp = ctypes.POINTER(ctypes.c_char_p)

# name -> (restype, argtypes)
SIGNATURES = {
    "Rdistances": (None, [c_fp, c_ip, c_ip, c_fp, c_cpp, c_fp]),
    "Rdistances_packed": (None, [c_fp, c_ip, c_ip, c_fp, c_cpp, c_fp]),
    "Rdistances_packed_prec": (None, [c_fp, c_ip, c_ip, c_fp, c_cpp, c_fp, c_cpp]),
    "Rdistclust": (None, [c_cpp, c_cpp, c_fp, c_ip, c_ip, c_ip, c_ip, c_fp]),
    "cuCompile": (None, [c_ip, c_cpp, c_cpp]),
    "unloadPackage": (None, []),
    "rsetDevice": (None, [c_ip]),
    "rgetDevice": (None, [c_ip]),
    "gt_last_error": (ctypes.c_char_p, []),
    "gt_version": (c_int, []),
    "gt_method_from_name": (c_int, [ctypes.c_char_p]),
    "gt_dist_workspace_size": (c_int, [c_i64, c_i64, c_int, c_int, ctypes.POINTER(c_size)]),
    "gt_dist": (c_int, [c_p, c_i64, c_i64, c_i64, c_int, c_float, c_int, c_int, c_p, c_i64, c_p]),
    "gt_dist_ws": (c_int, [c_p, c_i64, c_i64, c_i64, c_int, c_float, c_int, c_int, c_p, c_i64,
                           c_p, c_size, c_p]),
    "gt_dist_noalloc": (c_int, [c_p, c_i64, c_i64, c_i64, c_int, c_float, c_int, c_int, c_p, c_i64,
                                c_p]),
    "gt_dist_cross": (c_int, [c_p, c_i64, c_i64, c_p, c_i64, c_i64, c_i64, c_int, c_float, c_p,
                              c_i64, c_p]),
    "gt_dist_host": (c_int, [c_p, c_i64, c_i64, c_int, c_float, c_int, c_int, c_p]),
    "gt_dist_device_full": (c_int, [c_p, c_i64, c_i64, c_int, c_float, ctypes.POINTER(c_p),
                                    ctypes.POINTER(c_size)]),
    "gt_free": (c_int, [c_p]),
    "gt_dist_backward": (c_int, [c_p, c_i64, c_i64, c_i64, c_int, c_float, c_p, c_p, c_i64, c_p]),
    "gt_mds_stress": (c_int, [c_p, c_i64, c_i64, c_i64, c_p, c_p, c_i64, c_p, c_p]),
    "gt_dist_batched": (c_int, [c_p, c_p, c_p, c_p, c_i64, c_p, c_p, c_int, c_int, c_p, c_i64,
                                c_p]),
    "gt_dist_batched_schedule": (c_int, [c_p, c_p, c_i64, c_int, c_p, c_i64,
                                         ctypes.POINTER(c_i64)]),
    "gt_batched_tile_count": (c_int, []),
    "Rhcluster": (None, [c_fp, c_ip, c_ip, c_ip, c_fp, c_cpp]),
    "gt_hclust_method_from_name": (c_int, [ctypes.c_char_p]),
    "gt_hclust_workspace_size": (c_int, [c_i64, ctypes.POINTER(c_size)]),
    "gt_hclust": (c_int, [c_p, c_i64, c_i64, c_int, c_float, c_float, c_p, c_p, c_p, c_p, c_size,
                          c_p]),
    "gt_hclust_state": (c_int, [c_p, c_i64, c_i64, c_int, c_float, c_float, c_p, c_p, c_p, c_p,
                                c_size, c_int, c_p]),
    "gt_hclust_host": (c_int, [c_p, c_i64, c_i64, c_int, c_float, c_float, c_p, c_p, c_p]),
    "gt_hclust_format": (c_int, [c_i64, c_p, c_p, c_p, c_p]),
    "rpmcc": (None, [c_ip, c_fp, c_ip, c_fp, c_ip, c_ip, c_fp, c_fp, c_fp]),
    "rtestT": (None, [c_fp, c_fp, c_ip, c_fp]),
    "gt_pmcc_workspace_size": (c_int, [c_int, c_int, c_i64, c_i64, c_i64, c_int,
                                       ctypes.POINTER(c_size)]),
    "gt_pmcc": (c_int, [c_int, c_int, c_p, c_i64, c_i64, c_p, c_i64, c_i64, c_i64, c_p, c_p, c_p,
                        c_i64, c_p, c_size, c_p]),
    "gt_pmcc_host": (c_int, [c_int, c_int, c_p, c_i64, c_p, c_i64, c_i64, c_p, c_p, c_p]),
    "gt_ttest": (c_int, [c_p, c_p, c_i64, c_p, c_p]),
    "RgpuKendall": (None, [c_fp, c_ip, c_fp, c_ip, c_ip, ctypes.POINTER(ctypes.c_double)]),
    "gt_kendall_workspace_size": (c_int, [c_i64, c_i64, c_i64, c_int, ctypes.POINTER(c_size)]),
    "gt_kendall": (c_int, [c_p, c_i64, c_i64, c_p, c_i64, c_i64, c_i64, c_p, c_i64, c_p, c_size,
                           c_p]),
    "gt_kendall_host": (c_int, [c_p, c_i64, c_p, c_i64, c_i64, c_p]),
}


def lib():
    """Load the library once (raises LibraryNotBuilt when it is absent)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise LibraryNotBuilt(
                    "libgputools_amd.so is not built; run `python -m gputools_amd.build` "
                    "(there is no CPU fallback)")
            # torch bundles its own HIP runtime (SONAME libamdhip64.so.7); load
            # it first so that the library binds to the same runtime instance
            # that owns torch's device memory and streams.
            try:
                import torch  # noqa: F401
            except ImportError:  # pragma: no cover
                pass
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def last_error() -> str:
    return lib().gt_last_error().decode()


def check(status: int) -> None:
    if status != GT_OK:
        raise GputoolsError(status, last_error())
