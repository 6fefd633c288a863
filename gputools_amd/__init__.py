"""gputools_amd -- MI355X-native pairwise distances (gputools' gpuDist path),
the clustering stage of gpuDistClust / gpuHclust built on them, and gpuCor's
all-pairs correlations.

The compute is in libgputools_amd.so (hand-written gfx950 HIP kernels behind a
C-ABI, include/gputools_amd.h); this package is the host-side mirror of the
reference's R interface plus torch-tensor device entry points.
"""
from ._lib import GputoolsError, LibraryNotBuilt, lib  # noqa: F401
from .dist import (Dist, Rdistances, Rdistances_packed, Rdistances_packed_prec, dist, dist_cross, dist_device_full,  # noqa: F401
                   free_device, gpuDist, match_method, pmatch, workspace_size)
from .autograd import dist_backward, mds_stress, pdist  # noqa: F401
from .hclust import HClust, gpuDistClust, gpuHclust, hclust, hclust_format  # noqa: F401
from .cor import gpuCor, gpuTtest, kendall, pmcc, ttest  # noqa: F401
from .device import chooseGpu, getGpuId  # noqa: F401

__all__ = ["Dist", "GputoolsError", "LibraryNotBuilt", "Rdistances", "Rdistances_packed", "Rdistances_packed_prec", "dist", "dist_cross",
           "dist_device_full", "free_device", "gpuDist", "lib", "match_method", "pmatch",
           "workspace_size", "dist_backward", "mds_stress", "pdist", "HClust", "gpuDistClust",
           "gpuHclust", "hclust", "hclust_format", "gpuCor", "gpuTtest", "kendall", "pmcc", "ttest",
           "chooseGpu", "getGpuId"]
