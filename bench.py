"""Benchmark: all-pairs euclidean distances, n=65536 points x d=512 features.

One step = one gt_dist call over the whole point set (inputs resident in HBM,
packed FP32 lower-triangle output of n(n-1)/2 = 2,147,450,880 distances in
HBM).  The headline precision (default: BF16 three-term split, six MFMA
products per pair of terms' chunks, exact FP32 recompute of near-coincident
pairs) and the plain FP32 VALU kernel are both timed; value is the headline
path's pair-distances/s.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision bf16x3]

Prints one JSON line.  The input is the 64-centre Gaussian mixture (spread
1e-3) that stresses cancellation in the Gram form.  The cpu_baseline leg runs
the FP64 CPU oracle on every pair among a seeded sample of rows, reports its
throughput, and the same pairs give the max relative error of both GPU paths.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

N_POINTS = 65536
N_FEATURES = 512
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"bf16x3": 2500.0, "fp16x3": 2500.0, "fp32": 157.3}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="bf16x3", choices=["bf16x3", "fp16x3", "fp32"])
    ap.add_argument("--n", type=int, default=N_POINTS)
    ap.add_argument("--k", type=int, default=N_FEATURES)
    ap.add_argument("--fp32-steps", type=int, default=None,
                    help="timed steps of the plain FP32 comparison (default: --steps)")
    ap.add_argument("--sample-rows", type=int, default=11586,
                    help="rows whose all pairs form the CPU-baseline / error sample")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-paths", action="store_true",
                    help="skip the secondary paths (other methods, batched, backward, MDS, ...)")
    ap.add_argument("--paths", default=None,
                    help="comma list of secondary path groups to time (default: all of "
                         "methods,binary_large,uniform,batched,backward,small,hclust,pmcc,"
                         "kendall)")
    ap.add_argument("--paths-only", action="store_true",
                    help="time only the selected secondary paths (no headline line; for "
                         "profiling them)")
    ap.add_argument("--no-host-path", action="store_true",
                    help="skip timing the host-buffer entry point (PCIe-inclusive; never `value`)")
    ap.add_argument("--calibrate-peaks", action="store_true",
                    help="(default) measure sustained FP32 FMA, integer AND/popcount, BF16 MFMA, "
                         "copy and write-only fill rates with probe kernels (libgt_calib.so), "
                         "after the timed steps, and add them as measured_peaks")
    ap.add_argument("--no-peaks", action="store_true", help="skip the measured_peaks probes")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 2 GiB device-to-device copy first (reference kernel of known "
                         "size for scaling FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--exhaustive", action="store_true",
                    help="also compare every one of the n(n-1)/2 distances of each timed "
                         "precision with the FP64 oracle (streamed on the host cores; minutes)")
    ap.add_argument("--precisions", default=None,
                    help="comma list of paths to time (default: the headline and fp32)")
    return ap.parse_args()


def time_path(lib, x, out, ws, ws_bytes, precision, steps, warmup, torch):
    """W untimed steps, then K steps between two HIP events recorded on the
    stream the kernels are launched on (the mean per step).  An event between
    consecutive steps also gives each step's time, for the median."""
    from gputools_amd import _lib
    n, k = x.shape
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream or None
    prec = _lib.PRECISIONS[precision]

    def step():
        s = lib.gt_dist_ws(x.data_ptr(), n, k, x.stride(0), _lib.EUCLIDEAN, 2.0, prec,
                           _lib.LAYOUT_PACKED, out.data_ptr(), n,
                           ws.data_ptr() if ws is not None else None, ws_bytes, sp)
        if s != 0:
            raise RuntimeError(_lib.last_error())

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    ev[0].record(stream)
    for i in range(steps):
        step()
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    per_step = [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]
    return ev[0].elapsed_time(ev[steps]) / steps, float(np.median(per_step))


def sampled_pairs(n, rows):
    """Packed indices of every pair among the (sorted) sample rows."""
    r = np.sort(rows)
    jj, ii = np.triu_indices(r.size, 1)      # local j < i
    gi, gj = r[ii], r[jj]
    idx = gj * n - gj * (gj + 1) // 2 + (gi - gj - 1)
    return r, idx


def host_facts():
    """The CPU the baseline runs on: logical CPUs, the ones this process may
    use, the thread share (OMP_NUM_THREADS when set), the lscpu model, and
    whether R is installed (R's stats::dist is the named CPU baseline; it is
    timed only where Rscript exists, and nothing is installed)."""
    import shutil
    import subprocess
    facts = {"nproc": os.cpu_count()}
    try:
        facts["affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        facts["affinity"] = facts["nproc"]
    omp = os.environ.get("OMP_NUM_THREADS", "")
    facts["threads_available"] = int(omp) if omp.isdigit() and int(omp) > 0 else facts["affinity"]
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                facts["model"] = ln.split(":", 1)[1].strip()
            elif ln.startswith("Socket(s):"):
                facts["sockets"] = ln.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    rs = shutil.which("Rscript")
    if rs is None:
        facts["Rscript"] = "absent (Rscript --version: command not found)"
    else:
        try:
            r = subprocess.run([rs, "--version"], capture_output=True, text=True, timeout=30)
            facts["Rscript"] = (r.stdout + r.stderr).strip()[:200]
        except (OSError, subprocess.SubprocessError) as e:
            facts["Rscript"] = "present but failed: %s" % e
    return facts


def time_r_stats_dist(x):
    """Seconds R's stats::dist takes on x (rows = points), or None when R is
    not installed.  The points go to R through a binary file (readBin), so
    R's own RNG is never involved."""
    import shutil
    import subprocess
    import tempfile
    rs = shutil.which("Rscript")
    if rs is None:
        return None
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "x.bin")
        x.astype(np.float64).tofile(f)
        expr = ("x <- matrix(readBin('%s', 'double', n = %d), nrow = %d, byrow = TRUE); "
                "t <- system.time(d <- dist(x)); cat(t[['elapsed']])" % (f, x.size, x.shape[0]))
        r = subprocess.run([rs, "-e", expr], capture_output=True, text=True, timeout=600)
        try:
            return float(r.stdout.strip().split()[-1])
        except (ValueError, IndexError):
            return None


def _timed(fn, steps, warmup, torch, stream=None):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    st = stream or torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(st)
    for _ in range(steps):
        fn()
    b.record(st)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def _valu_roof(flops, ms):
    tf = flops / (ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": round(tf, 2), "peak": 157.3, "unit": "TFLOP/s",
            "frac": round(tf / 157.3, 4)}


def _roof(flops, hbm_bytes, ms):
    # the binding resource: FP32 VALU at 157.3 TFLOP/s or HBM at 8 TB/s,
    # whichever takes longer for this work
    if flops / 157.3e12 >= hbm_bytes / 8e12:
        return _valu_roof(flops, ms)
    gbs = hbm_bytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(gbs / 8000.0, 4)}


def path_methods(torch, gt, wl, steps, warmup, out):
    # all six methods, 8192 x 256 (FP32 VALU kernels; binary on bit rows)
    n, k = 8192, 256
    pairs = n * (n - 1) // 2
    meth = {}
    flops_per_feature = {"euclidean": 3, "maximum": 2, "manhattan": 2, "canberra": 4,
                         "minkowski": 3}
    for m in ("euclidean", "maximum", "manhattan", "canberra", "binary", "minkowski"):
        x = torch.from_numpy(wl.method_input(m, n, k, seed=2)).cuda()
        o = torch.empty(pairs, dtype=torch.float32, device="cuda")
        p = 3.0 if m == "minkowski" else 2.0
        ms = _timed(lambda: gt.dist(x, m, p, out=o), steps, warmup, torch)
        rec = {"ms": round(ms, 4), "pairs_per_s": round(pairs / (ms * 1e-3), 1)}
        if m == "binary":
            gbs = (4.0 * pairs + 4.0 * n * k) / (ms * 1e-3) / 1e9
            rec["roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0,
                               "unit": "GB/s", "frac": round(gbs / 8000.0, 4)}
        else:
            rec["roofline"] = _valu_roof(flops_per_feature[m] * k * pairs, ms)
        meth[m] = rec
    out["methods_8192x256_fp32"] = meth


def path_binary_large(torch, gt, wl, steps, warmup, out):
    # binary at 65536 x 256 (Bernoulli(0.1)): 8.59 GB of packed output, so
    # the HBM write of the result bounds it (1.07 ms at 8 TB/s)
    n, k = 65536, 256
    pairs = n * (n - 1) // 2
    x = torch.from_numpy(wl.bernoulli(n, k, seed=2)).cuda()
    o = torch.empty(pairs, dtype=torch.float32, device="cuda")
    ms = _timed(lambda: gt.dist(x, "binary", out=o), steps, warmup, torch)
    gbs = (4.0 * pairs + 4.0 * n * k) / (ms * 1e-3) / 1e9
    out["binary_65536x256"] = {
        "ms": round(ms, 4), "pairs_per_s": round(pairs / (ms * 1e-3), 1),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(gbs / 8000.0, 4)}}


def path_uniform(torch, gt, wl, steps, warmup, out):
    # euclidean on U[0,1) data, 65536 x 512: non-zero column means, which the
    # Gram form takes out (column centres) before splitting.  Error against
    # the FP64 oracle on every pair among 2000 seeded rows.
    from oracle import oracle as orc
    n, k = 65536, 512
    pairs = n * (n - 1) // 2
    xh = wl.uniform(n, k, seed=1)
    x = torch.from_numpy(xh).cuda()
    o = torch.empty(pairs, dtype=torch.float32, device="cuda")
    rows, idx = sampled_pairs(n, np.random.default_rng(7).choice(n, 2000, replace=False))
    ref = orc.dist(np.ascontiguousarray(xh[rows]), "euclidean")
    rec = {}
    for prec in ("bf16x3", "fp16x3", "fp32"):
        ms = _timed(lambda: gt.dist(x, "euclidean", precision=prec, out=o), steps, warmup, torch)
        got = o[torch.from_numpy(idx).cuda()].cpu().numpy().astype(np.float64)
        rec[prec] = {"ms": round(ms, 4), "pairs_per_s": round(pairs / (ms * 1e-3), 1),
                     "max_rel_err": float("%.3g" % np.max(np.abs(got - ref) / ref))}
    for prec in ("bf16x3", "fp16x3"):
        rec[prec]["speedup_vs_fp32"] = round(rec["fp32"]["ms"] / rec[prec]["ms"], 3)
    rec["error_sample_pairs"] = int(ref.size)
    out["uniform_65536x512"] = rec


def path_batched(torch, gt, wl, steps, warmup, out):
    # batched: 4096 sets, log-uniform sizes in [32, 2048], d in {16, 32, 64, 128}
    from gputools_amd.autotune import tune_batched
    from gputools_amd.batched import Batch
    sizes, dims = wl.batched_sets(4096, seed=5)
    rng = np.random.default_rng(55)
    xb = torch.from_numpy(rng.standard_normal(int((sizes * dims).sum()),
                                              dtype=np.float32)).cuda()
    bat = {}
    for m, fpf in (("euclidean", 3), ("manhattan", 2)):
        b = Batch(x=xb, sizes=sizes, dims=dims, method=m)
        tile = tune_batched(b, reps=3)
        g = b.capture(tile)
        ms_graph = _timed(g.replay, steps, warmup, torch)
        ms_launch = _timed(lambda: b.run(tile), steps, warmup, torch)
        ms_per_set = _timed(b.run_per_set, 1, 1, torch)
        flops = float(fpf * (sizes * (sizes - 1) // 2 * dims).sum())
        bat[m] = {"tile": tile, "ms_graph_replay": round(ms_graph, 4),
                  "ms_one_launch": round(ms_launch, 4), "ms_per_set_launches": round(ms_per_set, 4),
                  "speedup_vs_per_set": round(ms_per_set / ms_graph, 2),
                  "pairs_per_s": round(b.pairs / (ms_graph * 1e-3), 1),
                  "roofline": _valu_roof(flops, ms_graph)}
    out["batched_4096_sets"] = bat


def path_backward(torch, gt, wl, steps, warmup, out):
    # backward and MDS stress: delta from 16384 x 128 points
    n, k = 16384, 128
    pairs = n * (n - 1) // 2
    x = torch.from_numpy(wl.normal(n, k, seed=4)).cuda()
    delta = gt.dist(x)
    grad = torch.ones_like(delta)
    ms_bw = _timed(lambda: gt.dist_backward(x, grad, "euclidean"), steps, warmup, torch)
    # algorithmic work per pair: about 6k flop (the difference, its square and
    # sum to recompute d, and the weighted difference added to both rows' dX);
    # D is recomputed, not read.  The kernel's two-owner sweep does more.
    out["backward_euclidean_16384x128"] = {
        "ms": round(ms_bw, 4), "algorithmic_flops": 6.0 * k * pairs,
        "roofline": _valu_roof(6.0 * pairs * k, ms_bw)}
    mds = {}
    for q in (2, 32):
        y = torch.from_numpy(wl.normal(n, q, seed=44)).cuda()

        def step():
            yy = y.detach().requires_grad_(True)
            gt.mds_stress(yy, delta).backward()
        ms = _timed(step, steps, warmup, torch)
        # algorithmic work per pair: about 6q flop (as the backward) and one
        # 4-byte read of delta
        mds["q=%d" % q] = {"ms_forward_backward": round(ms, 4),
                           "algorithmic_flops": 6.0 * q * pairs, "algorithmic_bytes": 4.0 * pairs,
                           "roofline": _roof(6.0 * pairs * q, 4.0 * pairs, ms)}
    out["mds_stress_16384"] = mds


def path_small(torch, gt, wl, steps, warmup, out):
    # the reference's small case: 1000 x 64 uniform, euclidean, latency per call
    x = torch.from_numpy(wl.uniform(1000, 64, seed=1)).cuda()
    o = torch.empty(1000 * 999 // 2, dtype=torch.float32, device="cuda")
    out["uniform_1000x64_euclidean_ms"] = round(
        _timed(lambda: gt.dist(x, "euclidean", out=o), 20, 3, torch), 4)


def path_hclust(torch, gt, wl, steps, warmup, out):
    # clustering stage (gpuDistClust's second half) on the 8192 x 256 N(0,1)
    # euclidean matrix: n - 1 sequential merge rounds on one workgroup, so the
    # figure of merit is time per round.  The reference's loop rescans the
    # lower triangle every round: n^3/6 FP32 reads, whose HBM time alone is
    # given as the reference's floor.
    # Our kernel reads O(n^2) in all, so it is bound by the latency of its
    # n - 1 dependent rounds (argmin, update, repair, rescan, each ending in a
    # barrier), not by bandwidth.  The floor per round is measured with the
    # same kernel at n = 64, where every phase has almost nothing to do.
    def run_rounds(d, m):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        gt.hclust(d, m)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    small = gt.dist(torch.from_numpy(wl.normal(64, 256, seed=2)).cuda(), "euclidean", layout="full")
    floor = {}
    for m in ("single", "complete", "average", "ward"):
        run_rounds(small.clone(), m)
        floor[m] = min(run_rounds(small.clone(), m) for _ in range(3)) * 1e3 / 63
    n = 8192
    x = torch.from_numpy(wl.normal(n, 256, seed=2)).cuda()
    d0 = gt.dist(x, "euclidean", layout="full")
    hc = {"reference_floor_ms": round(n ** 3 / 6 * 4 / 8e12 * 1e3, 2)}
    for m in ("single", "complete", "average", "ward"):
        ms = run_rounds(d0.clone(), m)
        us = 1e3 * ms / (n - 1)
        hc[m] = {"ms": round(ms, 2), "us_per_round": round(us, 2),
                 "roofline": {"bound": "latency", "achieved": round(us, 2),
                              "floor": round(floor[m], 2), "unit": "us/round",
                              "frac": round(floor[m] / us, 4),
                              "floor_from": "the same kernel at n = 64"}}
    out["hclust_8192"] = hc


def path_pmcc(torch, gt, wl, steps, warmup, out):
    # gpuCor(x), Pearson: 8192 variables x 1024 samples N(0,1), the symmetric
    # self case (half the tiles computed, every value stored twice).  Each
    # pair gets pairs, corr and t-score: 12 B of output, which bounds the
    # "everything" path; its algorithmic work is one 1024-long dot product
    # per unique pair.  pairwise.complete.obs (10% NaN) is per-pair VALU
    # work: one pass of 6 fused products per sample (count, two shifted sums,
    # three second moments), 12 flops.
    nv, size = 8192, 1024
    a = torch.from_numpy(wl.normal(nv, size, seed=6)).cuda()
    outs = tuple(torch.empty((nv, nv), dtype=torch.float32, device="cuda") for _ in range(3))
    wbytes = 12.0 * nv * nv
    cor = {}
    for prec in ("bf16x3", "fp16x3", "fp32"):
        ms = _timed(lambda: gt.pmcc(a, precision=prec, out=outs), steps, warmup, torch)
        gbs = wbytes / (ms * 1e-3) / 1e9
        cor["everything_" + prec] = {
            "ms": round(ms, 4), "pairs_per_s": round(nv * nv / (ms * 1e-3), 1),
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0,
                         "unit": "GB/s", "frac": round(gbs / 8000.0, 4)}}
    rng = np.random.default_rng(66)
    an = a.clone()
    an[torch.from_numpy(rng.random((nv, size)) < 0.1).cuda()] = float("nan")
    ms = _timed(lambda: gt.pmcc(an, use="pairwise.complete.obs", precision="fp32", out=outs),
                steps, warmup, torch)
    cor["pairwise_complete_fp32"] = {
        "ms": round(ms, 4), "pairs_per_s": round(nv * nv / (ms * 1e-3), 1),
        "roofline": _valu_roof(12.0 * size * nv * (nv + 1) / 2, ms)}
    out["pmcc_self_8192x1024"] = cor


def path_kendall(torch, gt, wl, steps, warmup, out):
    # gpuCor(x, method = "kendall"): 1024 variables x 2048 samples N(0,1),
    # self case.  Algorithmic work of the count for variables without ties or
    # NaN (this workload): per unique pair of variables and 32 sample pairs,
    # one XOR and one popcount-accumulate (32-bit integer VALU lane
    # operations; 256 CUs x 64 lanes x 2.4 GHz = 39.3 T/s).
    nv, size = 1024, 2048
    a = torch.from_numpy(wl.normal(nv, size, seed=8)).cuda()
    tau = torch.empty((nv, nv), dtype=torch.float64, device="cuda")
    ms = _timed(lambda: gt.kendall(a, out=tau), steps, warmup, torch)
    q, r = divmod(size - 1, 32)
    words = 16 * q * (q + 1) + r * (q + 1)
    ops = 2.0 * words * nv * (nv + 1) / 2
    tops = ops / (ms * 1e-3) / 1e12
    out["kendall_self_1024x2048"] = {
        "ms": round(ms, 4), "pairs_per_s": round(nv * nv / (ms * 1e-3), 1),
        "sample_pairs_per_s": round(nv * (nv + 1) / 2 * size * (size - 1) / 2 / (ms * 1e-3), 1),
        "roofline": {"bound": "valu", "achieved": round(tops, 2), "peak": 39.3,
                     "unit": "T int32 lane-ops/s", "frac": round(tops / 39.3, 4)}}


PATHS = {"methods": path_methods, "binary_large": path_binary_large, "uniform": path_uniform,
         "batched": path_batched, "backward": path_backward,
         "small": path_small, "hclust": path_hclust, "pmcc": path_pmcc, "kendall": path_kendall}


def secondary_paths(torch, gt, wl, select=None, steps=3, warmup=1):
    """Timed secondary paths, each with its algorithmic work and roofline."""
    out = {}
    for name, fn in PATHS.items():
        if select is None or name in select:
            print("bench: path %s" % name, file=sys.stderr, flush=True)
            fn(torch, gt, wl, steps, warmup, out)
    return out


def measure_peaks(torch):
    """Sustained rates of register-only probe loops (gputools_amd/calib/
    gt_calib.hip) on this GPU: context for the published peaks the roofline
    fractions use.  2048 workgroups of 256 threads (8 per CU)."""
    path = os.path.join(ROOT, "gputools_amd", "libgt_calib.so")
    if not os.path.exists(path):
        return {"skipped": "libgt_calib.so not built"}
    L = ctypes.CDLL(path)
    vp, cf = ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)
    for name in ("gt_calib_fma", "gt_calib_popc", "gt_calib_mfma"):
        getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, cf]
        getattr(L, name).restype = ctypes.c_int
    L.gt_calib_copy.argtypes = [vp, vp, ctypes.c_int64, vp, cf]
    L.gt_calib_copy.restype = ctypes.c_int
    L.gt_calib_fill.argtypes = [vp, ctypes.c_int64, vp, cf]
    L.gt_calib_fill.restype = ctypes.c_int
    stream = torch.cuda.current_stream().cuda_stream or None
    blocks, threads = 2048, 256
    buf = torch.empty(blocks * threads, dtype=torch.float32, device="cuda")
    ms = ctypes.c_float(0.0)

    def run(fn, *a):
        for _ in range(2):          # the first launch loads the code object
            if fn(*a, stream, ctypes.byref(ms)) != 0:
                raise RuntimeError("calibration probe failed")
        return ms.value

    out = {"probe": "gputools_amd/calib/gt_calib.hip, %d x %d threads" % (blocks, threads)}
    it = 20000
    t = run(L.gt_calib_fma, blocks, it, buf.data_ptr())
    lanes = blocks * threads * it * 16.0
    out["fp32_fma"] = {"ms": round(t, 3), "tflops": round(2 * lanes / (t * 1e-3) / 1e12, 1),
                       "t_lane_instr_per_s": round(lanes / (t * 1e-3) / 1e12, 1),
                       "what": "16 independent non-packed v_fma_f32 chains per lane"}
    t = run(L.gt_calib_popc, blocks, it, buf.data_ptr())
    ops = blocks * threads * it * 32.0
    out["int32_and_popcount"] = {"ms": round(t, 3),
                                 "t_lane_ops_per_s": round(ops / (t * 1e-3) / 1e12, 1),
                                 "what": "16 independent v_and_b32 + v_bcnt_u32_b32 chains"}
    it = 200
    t = run(L.gt_calib_mfma, blocks, it, buf.data_ptr())
    flops = blocks * (threads // 64) * it * 8 * 32768.0
    out["bf16_mfma"] = {"ms": round(t, 3), "pflops": round(flops / (t * 1e-3) / 1e15, 3),
                        "what": "8 independent v_mfma_f32_32x32x16_bf16 chains per wave"}
    del buf
    n4 = (1 << 30) // 16
    a = torch.ones(n4 * 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    t = run(L.gt_calib_copy, a.data_ptr(), b.data_ptr(), n4)
    out["copy"] = {"ms": round(t, 3), "gbs_read_plus_write": round(2 * 16.0 * n4 / (t * 1e-3) / 1e9, 1),
                   "what": "1 GiB float4 device-to-device copy"}
    del a, b
    n4 = (8 << 30) // 16
    a = torch.empty(n4 * 4, dtype=torch.float32, device="cuda")
    t = run(L.gt_calib_fill, a.data_ptr(), n4)
    out["fill"] = {"ms": round(t, 3), "gbs_write": round(16.0 * n4 / (t * 1e-3) / 1e9, 1),
                   "what": "8 GiB float4 write-only fill (the size of the packed 65536-point output)"}
    del a
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    import gputools_amd as gt
    from gputools_amd import workloads as wl

    lib = gt.lib()
    select = None if args.paths is None else set(p for p in args.paths.split(",") if p)
    if select is not None and not select <= set(PATHS):
        print("bench.py: unknown path group(s) %s" % sorted(select - set(PATHS)), file=sys.stderr)
        sys.exit(2)
    if args.paths_only:
        print(json.dumps({"paths": secondary_paths(torch, gt, wl, select, args.steps,
                                                    args.warmup)}))
        return
    if args.calibrate_copy:
        a = torch.ones(512 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)       # reads 2 GiB, writes 2 GiB
        torch.cuda.synchronize()
        del a, b
    n, k = args.n, args.k
    npairs = n * (n - 1) // 2
    print("bench: generating %dx%d mixture" % (n, k), file=sys.stderr, flush=True)
    xh = wl.mixture(n, k, seed=3)
    x = torch.from_numpy(xh).cuda()
    out = torch.empty(npairs, dtype=torch.float32, device="cuda")

    results = {}
    if args.precisions:
        order = [p for p in args.precisions.split(",") if p]
        if args.precision not in order:
            order.insert(0, args.precision)
    else:
        order = [args.precision, "fp32"] if args.precision != "fp32" else ["fp32"]
    for precision in order:
        ws_bytes = gt.workspace_size(n, k, "euclidean", precision)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda") if ws_bytes else None
        steps = args.steps if precision == args.precision else (args.fp32_steps or args.steps)
        print("bench: timing %s, %d warmup + %d steps" % (precision, args.warmup, steps),
              file=sys.stderr, flush=True)
        ms, med = time_path(lib, x, out, ws, ws_bytes, precision, steps, args.warmup, torch)
        rec = {"ms_per_step": ms, "median_ms": med, "pairs_per_s": npairs / (ms * 1e-3)}
        if not args.no_cpu_baseline:
            rng = np.random.default_rng(12345)
            rows = rng.choice(n, size=min(args.sample_rows, n), replace=False)
            rec["_rows"], idx = sampled_pairs(n, rows)
            rec["_sample"] = out[torch.from_numpy(idx).cuda()].cpu().numpy()
        results[precision] = rec
        print("bench: %s %.3f ms/step" % (precision, ms), file=sys.stderr, flush=True)

    head = results[args.precision]
    ms = head["ms_per_step"]
    # Algorithmic work per launch: packed FP32 output (4 B/pair) plus one read
    # of the n x k FP32 input; 2k flops/pair for the Gram form, 3k for the
    # difference form.
    bytes_per_launch = 4.0 * npairs + 4.0 * n * k
    flops_per_launch = (2.0 if args.precision != "fp32" else 3.0) * k * npairs
    gbs = bytes_per_launch / (ms * 1e-3) / 1e9
    tflops = flops_per_launch / (ms * 1e-3) / 1e12
    t_hbm = bytes_per_launch / (HBM_PEAK_GBS * 1e9)
    t_cmp = flops_per_launch / (PEAK_TFLOPS[args.precision] * 1e12)
    if t_hbm >= t_cmp:
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4)}
    else:
        roof = {"bound": "mfma" if args.precision != "fp32" else "valu",
                "achieved": round(tflops, 2), "peak": PEAK_TFLOPS[args.precision],
                "unit": "TFLOP/s", "frac": round(tflops / PEAK_TFLOPS[args.precision], 4)}
    roof["traffic"] = None
    # HBM bytes per launch of the dominant kernel, from the FETCH_SIZE /
    # WRITE_SIZE counter runs recorded under profiles/ (scaled by the same
    # counters' reading of a 2 GiB copy); null when no record matches.
    tpath = os.path.join(ROOT, "profiles", "traffic_mixture_65536x512.json")
    try:
        with open(tpath) as f:
            rec = json.load(f).get(args.precision)
        if rec and n == N_POINTS and k == N_FEATURES:
            roof["traffic"] = rec["hbm_bytes_per_launch"]
            roof["traffic_source"] = "profiles/traffic_mixture_65536x512.json"
            roof["traffic_over_algorithmic"] = round(rec["hbm_bytes_per_launch"] /
                                                     bytes_per_launch, 2)
    except (OSError, ValueError, KeyError):
        pass
    roof["kernel_ms"] = round(ms, 4)
    roof["algorithmic_bytes"] = bytes_per_launch
    roof["algorithmic_flops"] = flops_per_launch

    line = {
        "metric": "Euclidean pair-dists/s n=65536 d=512: BF16-split vs FP32, at <=1e-5 rel err",
        "value": round(npairs / (ms * 1e-3), 1),
        "unit": "pair-dists/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "median_ms_per_step": round(head["median_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x3": "bf16", "fp16x3": "fp16", "fp32": "fp32"}[args.precision],
        "data": "synthetic: 64-centre Gaussian mixture, spread 1e-3, numpy seed 3",
        "config": {"workload": "gpuDist euclidean all pairs, n=%d x d=%d, packed dist output"
                               % (n, k), "n": n, "d": k, "method": "euclidean",
                   "precision": args.precision, "pairs": npairs},
        "roofline": roof,
    }
    if args.calibrate_peaks or not args.no_peaks:
        print("bench: calibration probes", file=sys.stderr, flush=True)
        try:
            line["measured_peaks"] = measure_peaks(torch)
        except (RuntimeError, OSError) as e:   # torch's out-of-memory error is a RuntimeError
            line["measured_peaks"] = {"skipped": str(e)[:200]}
            torch.cuda.empty_cache()
    if args.precision != "fp32" and "fp32" in results:
        f = results["fp32"]
        line["fp32_plain"] = {"ms_per_step": round(f["ms_per_step"], 4),
                              "median_ms_per_step": round(f["median_ms"], 4),
                              "pairs_per_s": round(f["pairs_per_s"], 1)}
        line["speedup_vs_fp32"] = round(f["ms_per_step"] / ms, 3)
        line["speedup_vs_fp32_median"] = round(f["median_ms"] / head["median_ms"], 3)
    for other, rec in results.items():
        if other not in (args.precision, "fp32"):
            line.setdefault("other_precisions", {})[other] = {
                "ms_per_step": round(rec["ms_per_step"], 4),
                "pairs_per_s": round(rec["pairs_per_s"], 1),
                "speedup_vs_fp32": (round(results["fp32"]["ms_per_step"] / rec["ms_per_step"], 3)
                                    if "fp32" in results else None)}

    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        rows = head["_rows"]
        xs = np.ascontiguousarray(xh[rows])
        host = host_facts()
        threads = args.cpu_threads or host["threads_available"]
        print("bench: cpu baseline, oracle on %d rows, %d threads" % (rows.size, threads),
              file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        ref = orc.dist(xs, "euclidean", nthreads=threads)
        dt = time.perf_counter() - t0
        # one core, on a smaller sample of the same rows (its pairs are a
        # subset of the sample above)
        r1 = min(rows.size, 3000)
        t1 = time.perf_counter()
        one = orc.dist(np.ascontiguousarray(xs[:r1]), "euclidean", nthreads=1)
        dt1 = time.perf_counter() - t1
        line["cpu_baseline"] = {"value": round(ref.size / dt, 1), "unit": "pair-dists/s",
                                "cores": threads, "kind": "port",
                                "sample": "all %d pairs among %d seeded rows of the %dx%d "
                                          "mixture, FP64 oracle (the restatement of R "
                                          "stats::dist), %.2f s on %d threads"
                                          % (ref.size, rows.size, n, k, dt, threads),
                                "one_core": {"value": round(one.size / dt1, 1), "cores": 1,
                                             "sample": "all %d pairs among the first %d of "
                                                       "those rows, %.2f s" % (one.size, r1, dt1)},
                                "host": host}
        rtime = time_r_stats_dist(np.ascontiguousarray(xs[:r1]))
        if rtime is not None:
            line["cpu_baseline"]["r_stats_dist"] = {
                "value": round(one.size / rtime, 1), "cores": 1,
                "sample": "R stats::dist on the same %d rows (FP64), %.2f s" % (r1, rtime)}
        errs = {}
        for precision, rec in results.items():
            g = rec["_sample"].astype(np.float64)
            errs[precision] = float(np.max(np.abs(g - ref) / ref))
        line["max_rel_err"] = errs[args.precision]
        line["max_rel_err_by_path"] = {p: float("%.3g" % e) for p, e in errs.items()}
        line["error_sample_pairs"] = int(ref.size)

    if not args.no_host_path:
        # The host-buffer boundary (gt_dist_host, the counterpart of distance()):
        # H2D of the points, the kernels, D2H of the packed result into pageable
        # host memory, device allocations included.  Never `value`.
        from gputools_amd import _lib
        out_h = np.empty(npairs, dtype=np.float32)
        prec = _lib.PRECISIONS[args.precision]

        def host_call():
            s = lib.gt_dist_host(xh.ctypes.data, n, k, _lib.EUCLIDEAN, 2.0, prec,
                                 _lib.LAYOUT_PACKED, out_h.ctypes.data)
            if s != 0:
                raise RuntimeError(_lib.last_error())

        print("bench: host-buffer path, 1 warm-up + 1 timed call", file=sys.stderr, flush=True)
        host_call()                          # also first-touches the output pages
        t0 = time.perf_counter()
        host_call()
        hms = (time.perf_counter() - t0) * 1e3
        # the transfers alone, from and to pageable host memory that has been
        # touched before (the host path's own buffers), as torch copies
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        torch.from_numpy(xh).cuda()
        torch.cuda.synchronize()
        h2d = (time.perf_counter() - t0) * 1e3
        dst = torch.from_numpy(out_h)
        t0 = time.perf_counter()
        dst.copy_(out)
        d2h = (time.perf_counter() - t0) * 1e3
        line["host_path"] = {"ms": round(hms, 2), "pairs_per_s": round(npairs / (hms * 1e-3), 1),
                             "h2d_points_ms": round(h2d, 2),
                             "d2h_packed_result_ms": round(d2h, 2),
                             "d2h_gb_per_s": round(4.0 * npairs / (d2h * 1e-3) / 1e9, 2),
                             "what": "gt_dist_host: H2D of X, kernels, D2H of the %.2f GB packed "
                                     "result to pageable host memory, allocations included"
                                     % (4.0 * npairs / 1e9)}
        del out_h

    if args.exhaustive:
        # every pair of the workload against the FP64 oracle: the packed result
        # of one more call per precision, copied to the host and compared by
        # the oracle's streaming comparator on all host threads
        from oracle import oracle as orc
        threads = host_facts()["threads_available"]
        ex = {}
        for precision in results:
            ws_bytes = gt.workspace_size(n, k, "euclidean", precision)
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda") if ws_bytes else None
            time_path(lib, x, out, ws, ws_bytes, precision, 1, 0, torch)
            del ws
            got = out.cpu().numpy()
            print("bench: exhaustive check of %s, %d pairs, %d threads" % (precision, npairs, threads),
                  file=sys.stderr, flush=True)
            t0 = time.perf_counter()
            err, (wi, wj) = orc.compare_packed(xh, got, "euclidean", nthreads=threads)
            ex[precision] = {"max_rel_err": float("%.4g" % err), "worst_pair": [int(wi), int(wj)],
                             "pairs": npairs, "seconds": round(time.perf_counter() - t0, 1)}
            print("bench: %s max rel err %.3g at (%d, %d)" % (precision, err, wi, wj),
                  file=sys.stderr, flush=True)
            del got
        line["exhaustive_check"] = ex

    if not args.no_paths:
        print("bench: secondary paths", file=sys.stderr, flush=True)
        line["paths"] = secondary_paths(torch, gt, wl, select)

    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
