"""CPU-only checks of the C-ABI library and the host-side mirror.

No compute call is made here (there is no GPU in this suite): the library must
load, export every symbol include/gputools_amd.h declares, map method names as
getDistEnum does, and reject bad arguments before touching the device.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import gputools_amd as gt
from gputools_amd import _lib
from gputools_amd.dist import Dist, match_method, pmatch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gputools_amd.h")


def declared_functions():
    with open(HEADER) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"defined"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def test_library_loads_and_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert "Rdistances" in names and "gt_dist" in names and "rsetDevice" in names
    for name in names:
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, "ctypes signature missing for " + name


def test_exported_symbols_are_unmangled():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for name in declared_functions():
        assert re.search(r"\bT %s$" % name, out, flags=re.M), name


def test_method_enum_matches_getDistEnum():
    L = _lib.lib()
    table = {b"euclidean": 0, b"maximum": 1, b"manhattan": 2, b"canberra": 3, b"binary": 4,
             b"minkowski": 5}
    for name, val in table.items():
        assert L.gt_method_from_name(name) == val
    # unknown names fall back to euclidean, as getDistEnum does
    for name in (b"dot", b"", b"Euclidean", b"max"):
        assert L.gt_method_from_name(name) == 0
    assert L.gt_method_from_name(None) == 0


def test_version():
    assert _lib.lib().gt_version() >= 10000


def _status(fn, *args):
    return fn(*args), _lib.last_error()


def test_argument_validation_happens_before_the_device():
    L = _lib.lib()
    dummy = ctypes.c_void_p(16)   # never dereferenced: every call fails validation
    # unknown method
    s, msg = _status(L.gt_dist, dummy, 10, 4, 4, 9, 2.0, 0, 0, dummy, 10, None)
    assert s == _lib.GT_ERR_INVALID and "method" in msg
    # minkowski with p <= 0 / inf
    for p in (0.0, -2.0, float("inf")):
        s, msg = _status(L.gt_dist, dummy, 10, 4, 4, 5, p, 0, 0, dummy, 10, None)
        assert s == _lib.GT_ERR_INVALID and "p" in msg
    # negative sizes, short leading dimension
    assert L.gt_dist(dummy, -1, 4, 4, 0, 2.0, 0, 0, dummy, 10, None) == _lib.GT_ERR_INVALID
    assert L.gt_dist(dummy, 10, 4, 3, 0, 2.0, 0, 0, dummy, 10, None) == _lib.GT_ERR_INVALID
    # full layout with ldd < n
    assert L.gt_dist(dummy, 10, 4, 4, 0, 2.0, 0, 1, dummy, 9, None) == _lib.GT_ERR_INVALID
    # unknown layout / precision
    assert L.gt_dist(dummy, 10, 4, 4, 0, 2.0, 0, 7, dummy, 10, None) == _lib.GT_ERR_INVALID
    assert L.gt_dist(dummy, 10, 4, 4, 0, 2.0, 9, 0, dummy, 10, None) == _lib.GT_ERR_INVALID
    # split precision is euclidean-only
    s, msg = _status(L.gt_dist, dummy, 10, 4, 4, 2, 2.0, 1, 0, dummy, 10, None)
    assert s == _lib.GT_ERR_UNSUPPORTED
    # null pointers
    assert L.gt_dist(None, 10, 4, 4, 0, 2.0, 0, 0, dummy, 10, None) == _lib.GT_ERR_INVALID
    assert L.gt_dist(dummy, 10, 4, 4, 0, 2.0, 0, 0, None, 10, None) == _lib.GT_ERR_INVALID
    # workspace too small for binary
    s, _ = _status(L.gt_dist_ws, dummy, 10, 40, 40, 4, 2.0, 0, 0, dummy, 10, None, 0, None)
    assert s == _lib.GT_ERR_WORKSPACE


def test_empty_inputs_are_no_ops():
    L = _lib.lib()
    # n < 2 packed: nothing to compute, no device call
    assert L.gt_dist(None, 0, 4, 4, 0, 2.0, 0, 0, None, 0, None) == _lib.GT_OK
    assert L.gt_dist(None, 1, 4, 4, 0, 2.0, 0, 0, None, 1, None) == _lib.GT_OK


def test_workspace_sizes():
    b = ctypes.c_size_t()
    L = _lib.lib()
    assert L.gt_dist_workspace_size(1000, 64, 0, 0, ctypes.byref(b)) == 0 and b.value == 0
    assert L.gt_dist_workspace_size(1000, 70, 4, 0, ctypes.byref(b)) == 0
    assert b.value >= 1000 * 3 * 4          # 70 features -> 3 words per row
    assert L.gt_dist_workspace_size(1000, 64, 0, 1, ctypes.byref(b)) == 0
    assert b.value >= 1000 * 4              # row norms
    assert L.gt_dist_workspace_size(1000, 64, 0, 0, None) == _lib.GT_ERR_INVALID


def test_cuCompile_is_a_noop_with_the_R_signature():
    L = _lib.lib()
    names = (ctypes.c_char_p * 1)(b"distance")
    srcs = (ctypes.c_char_p * 1)(b"__global__ void k(){}")
    L.cuCompile(ctypes.byref(ctypes.c_int(1)), names, srcs)
    L.unloadPackage()


def test_pmatch_semantics():
    table = ["euclidean", "maximum", "manhattan", "canberra", "binary", "minkowski"]
    assert pmatch("euclidean", table) == 0
    assert pmatch("max", table) == 1
    assert pmatch("man", table) == 2
    assert pmatch("m", table) is None        # ambiguous -> NA
    assert pmatch("", table) is None         # empty string matches nothing
    assert pmatch("canb", table) == 3
    assert pmatch("euclideans", table) is None


def test_match_method_follows_gpuDist():
    assert match_method("euclidian") == "euclidean"     # alias
    assert match_method("eu") == "euclidean"            # prefix of the alias
    assert match_method("mink") == "minkowski"
    assert match_method("binary") == "binary"
    for bad in ("m", "", "cosine", "Euclidean"):
        with pytest.raises(ValueError, match="invalid distance method"):
            match_method(bad)


def test_gpuDist_rejects_non_finite_input_like_dotC():
    x = np.ones((4, 3))
    x[1, 2] = np.nan
    with pytest.raises(ValueError, match="NA/NaN/Inf"):
        gt.gpuDist(x)
    x[1, 2] = np.inf
    with pytest.raises(ValueError, match="NA/NaN/Inf"):
        gt.gpuDist(x)


def test_dist_object_layout():
    n = 5
    vals = np.arange(n * (n - 1) // 2, dtype=np.float64) + 1
    d = Dist(vals, n, method="manhattan")
    m = d.as_matrix()
    assert np.all(m == m.T) and np.all(np.diag(m) == 0)
    # R order: column 0 holds pairs (1,0), (2,0), (3,0), (4,0)
    assert [d[i, 0] for i in range(1, 5)] == [1, 2, 3, 4]
    assert d[2, 1] == 5 and d[1, 2] == 5
    assert d.attrs["Diag"] is False and d.attrs["Upper"] is False and "p" not in d.attrs
    with pytest.raises(IndexError):
        d.index(3, 3)


def test_noalloc_rejects_oversized_requests_before_the_device():
    """gt_dist_noalloc: a request whose workspace exceeds the static scratch
    fails with GT_ERR_WORKSPACE, naming the limit, without any device call."""
    L = _lib.lib()
    dummy = ctypes.c_void_p(16)
    n, k = 65536, 512
    need = gt.workspace_size(n, k, "euclidean", "bf16x3")
    assert need > _lib.NOALLOC_SCRATCH_BYTES
    s, msg = _status(L.gt_dist_noalloc, dummy, n, k, k, 0, 2.0, _lib.PREC_BF16X3,
                     _lib.LAYOUT_PACKED, dummy, n, None)
    assert s == _lib.GT_ERR_WORKSPACE, msg
    assert str(_lib.NOALLOC_SCRATCH_BYTES) in msg and "gt_dist_ws" in msg
    # binary rows of 2^20 points x 4096 features: 129 words each, 541 MB
    s, msg = _status(L.gt_dist_noalloc, dummy, 1 << 20, 4096, 4096, _lib.BINARY, 2.0,
                     _lib.PREC_FP32, _lib.LAYOUT_PACKED, dummy, 1 << 20, None)
    assert s == _lib.GT_ERR_WORKSPACE, msg
    # ordinary argument errors come first, as for gt_dist
    s, msg = _status(L.gt_dist_noalloc, dummy, 10, 4, 4, 9, 2.0, 0, 0, dummy, 10, None)
    assert s == _lib.GT_ERR_INVALID and "method" in msg
