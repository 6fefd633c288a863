// gt_capi.cpp -- the C-ABI of libgputools_amd.so (declared in
// include/gputools_amd.h): argument checking, workspace handling and
// dispatch to the gfx950 kernels.
//
// Replaces gputools' host driver (src/distance.cpp:209-344), its .C glue
// (src/rinterface.cpp:79-85, 147-156, 172-214) and its NVRTC runtime
// (src/cudaUtils.cpp:117-218): kernels are precompiled, nothing is JIT-compiled
// or module-loaded per call, errors become status codes (or R errors when the
// library runs inside R) instead of longjmps that leak device buffers.
#include "../../include/gputools_amd.h"
#include "gt_internal.h"

#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <limits>
#include <mutex>
#include <vector>

// Workspace of gt_dist_noalloc: a static device array, one copy per GPU,
// held from module load until the process ends.  Its name is a host-side
// stand-in; host code uses only the device address that hipGetSymbolAddress
// returns for the current device (scratch_address below).
static __device__ __align__(256) unsigned char scratch[GT_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local char g_err[1024] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(e == hipErrorOutOfMemory ? GT_ERR_ALLOC : GT_ERR_HIP, "%s: %s", what,
                hipGetErrorString(e));
}

#define GT_HIP(call)                                                    \
    do {                                                                \
        hipError_t e_ = (call);                                         \
        if (e_ != hipSuccess) return hip_fail(e_, #call);               \
    } while (0)

// When the library is loaded into an R process, a failed .C entry point
// raises an R error (the convention of checkCudaError, src/cuseful.cpp:95-102);
// otherwise the message stays available through gt_last_error().  Called only
// after every resource of the entry point has been released, because Rf_error
// does not return.
void raise_in_r(int status)
{
    if (status == GT_OK) return;
    typedef void (*rf_error_t)(const char *, ...);
    static rf_error_t rf_error = (rf_error_t)dlsym(RTLD_DEFAULT, "Rf_error");
    if (rf_error) rf_error("%s", g_err);
}

bool package_loaded = false;

int check_method(int method, float p)
{
    if (method < GT_EUCLIDEAN || method > GT_MINKOWSKI)
        return fail(GT_ERR_INVALID, "unknown distance method %d", method);
    if (method == GT_MINKOWSKI && !(p > 0.0f && isfinite(p)))
        return fail(GT_ERR_INVALID, "invalid p %g for minkowski (must be finite and > 0)",
                    (double)p);
    return GT_OK;
}

// Per-pair operation of the FP32 path for a method (minkowski p = 1, 2, 3 are
// the manhattan, euclidean and cube specialisations).
int plain_op(int method, float p)
{
    switch (method) {
    case GT_EUCLIDEAN: return gt::OP_EUCLID;
    case GT_MAXIMUM: return gt::OP_MAXIMUM;
    case GT_MANHATTAN: return gt::OP_MANHATTAN;
    case GT_CANBERRA: return gt::OP_CANBERRA;
    case GT_MINKOWSKI:
        if (p == 1.0f) return gt::OP_MANHATTAN;
        if (p == 2.0f) return gt::OP_EUCLID;
        if (p == 3.0f) return gt::OP_MINK3;
        return gt::OP_MINKP;
    default: return -1;
    }
}

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

int workspace_bytes(int64_t n, int64_t k, int method, int precision, size_t *bytes)
{
    size_t b = 0;
    if (k > 0 && n > 1) {
        if (method == GT_BINARY)
            b = gt::binary_workspace_bytes(n, k);
        else if (precision != GT_PREC_FP32)
            b = gt::gram_workspace_bytes(n, k, precision);
    }
    *bytes = b;
    return GT_OK;
}

int check_dist_args(const float *x, int64_t n, int64_t k, int64_t ldx, int method, float p,
                    int precision, int layout, const float *out, int64_t ldd)
{
    int s = check_method(method, p);
    if (s != GT_OK) return s;
    if (n < 0 || k < 0) return fail(GT_ERR_INVALID, "negative size (n=%lld, k=%lld)",
                                    (long long)n, (long long)k);
    if (n > (int64_t)std::numeric_limits<int>::max() / 2)
        return fail(GT_ERR_INVALID, "n=%lld exceeds the supported maximum", (long long)n);
    if (ldx < k) return fail(GT_ERR_INVALID, "ldx=%lld < k=%lld", (long long)ldx, (long long)k);
    if (layout != GT_LAYOUT_PACKED && layout != GT_LAYOUT_FULL)
        return fail(GT_ERR_INVALID, "unknown layout %d", layout);
    if (layout == GT_LAYOUT_FULL && ldd < n)
        return fail(GT_ERR_INVALID, "ldd=%lld < n=%lld", (long long)ldd, (long long)n);
    if (precision < GT_PREC_FP32 || precision > GT_PREC_FP16X3)
        return fail(GT_ERR_INVALID, "unknown precision %d", precision);
    if (precision != GT_PREC_FP32 && method != GT_EUCLIDEAN)
        return fail(GT_ERR_UNSUPPORTED,
                    "split-precision Gram form exists for euclidean only (method %d)", method);
    // the Gram epilogue addresses a wave's 64 output rows with 32-bit offsets
    // (an output of that size could not be allocated anyway)
    if (precision != GT_PREC_FP32 && (n >= (int64_t(1) << 24) || (layout == GT_LAYOUT_FULL && ldd >= (int64_t(1) << 24))))
        return fail(GT_ERR_INVALID, "split-precision path supports n and ldd below 2^24");
    const bool has_out = (layout == GT_LAYOUT_PACKED) ? n > 1 : n > 0;
    if (has_out && out == nullptr) return fail(GT_ERR_INVALID, "null output pointer");
    if (has_out && k > 0 && x == nullptr) return fail(GT_ERR_INVALID, "null input pointer");
    return GT_OK;
}

int run_dist(const float *x, int64_t n, int64_t k, int64_t ldx, int method, float p,
             int precision, int layout, float *out, int64_t ldd, void *ws, hipStream_t s)
{
    const int glayout = (layout == GT_LAYOUT_PACKED) ? gt::OUT_PACKED : gt::OUT_FULL;
    if (layout == GT_LAYOUT_PACKED && n < 2) return GT_OK;
    if (n < 1) return GT_OK;
    if (k == 0) {
        // No usable feature: R returns NA for every pair.
        const float na = __builtin_nanf("");
        if (layout == GT_LAYOUT_PACKED)
            GT_HIP(gt::launch_fill(out, n * (n - 1) / 2, na, s));
        else
            GT_HIP(gt::launch_fill_full(out, n, ldd, na, s));
        return GT_OK;
    }
    if (method == GT_BINARY) {
        if (n == 1) {
            GT_HIP(gt::launch_fill_full(out, 1, ldd, 0.0f, s));
            return GT_OK;
        }
        GT_HIP(gt::launch_binary_self(glayout, x, n, k, ldx, ws, out, ldd, s));
        return GT_OK;
    }
    if (precision != GT_PREC_FP32) {
        GT_HIP(gt::launch_gram(precision, glayout, x, n, k, ldx, out, ldd, ws, s));
        return GT_OK;
    }
    GT_HIP(gt::launch_plain(plain_op(method, p), glayout, x, n, ldx, x, n, ldx, k, p, out, ldd, s));
    return GT_OK;
}

// Fault injection for tests of the failure paths: with GT_FAIL_ALLOC=N
// (N >= 1) in the environment, the Nth device allocation the library makes
// after the variable was last set (or changed) fails as if the device were
// out of memory.  The reference longjmps out of such a failure and leaks
// what it had allocated (src/distance.cpp:277-311, src/cuseful.cpp:95-102);
// here every entry point releases what it holds and returns GT_ERR_ALLOC.
std::mutex g_fi_mu;
char g_fi_value[32] = "";
long g_fi_seen = 0;

bool injected_alloc_failure()
{
    const char *v = getenv("GT_FAIL_ALLOC");
    std::lock_guard<std::mutex> lock(g_fi_mu);
    if (v == nullptr || *v == '\0') {
        g_fi_value[0] = '\0';
        return false;
    }
    if (strncmp(v, g_fi_value, sizeof(g_fi_value) - 1) != 0) {
        strncpy(g_fi_value, v, sizeof(g_fi_value) - 1);
        g_fi_value[sizeof(g_fi_value) - 1] = '\0';
        g_fi_seen = 0;
    }
    const long n = strtol(v, nullptr, 10);
    return n >= 1 && ++g_fi_seen == n;
}

hipError_t device_malloc(void **p, size_t bytes)
{
    if (injected_alloc_failure()) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

// Stream-ordered scratch: hipMallocAsync / hipFreeAsync on the caller's stream
// (safe against concurrent calls on other streams).  Falls back to a
// synchronous hipMalloc/hipFree pair only where the stream-ordered pool is
// not supported, and never while the stream is being captured into a graph
// (a synchronous allocation or synchronisation would invalidate the capture):
// callers that capture pass their own workspace (the *_ws entry points) or
// use gt_dist_noalloc.
struct Scratch {
    void *ptr = nullptr;
    bool async = false;
    hipStream_t stream = nullptr;

    int acquire(size_t bytes, hipStream_t s)
    {
        stream = s;
        if (bytes == 0) return GT_OK;
        hipError_t e = injected_alloc_failure() ? hipErrorOutOfMemory : hipMallocAsync(&ptr, bytes, s);
        if (e == hipSuccess) {
            async = true;
            return GT_OK;
        }
        (void)hipGetLastError();
        ptr = nullptr;
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cap) == hipSuccess && cap != hipStreamCaptureStatusNone)
            return fail(GT_ERR_WORKSPACE,
                        "stream-ordered workspace allocation failed during graph capture (%s); "
                        "pass a workspace (gt_dist_ws) or call gt_dist_noalloc",
                        hipGetErrorString(e));
        if (e != hipErrorNotSupported) return hip_fail(e, "hipMallocAsync(workspace)");
        e = device_malloc(&ptr, bytes);
        if (e != hipSuccess) {
            ptr = nullptr;
            return hip_fail(e, "hipMalloc(workspace)");
        }
        return GT_OK;
    }

    int release()
    {
        if (!ptr) return GT_OK;
        hipError_t e;
        if (async) {
            e = hipFreeAsync(ptr, stream);
        } else {
            e = hipStreamSynchronize(stream);
            hipError_t e2 = hipFree(ptr);
            if (e == hipSuccess) e = e2;
        }
        ptr = nullptr;
        return e == hipSuccess ? GT_OK : hip_fail(e, "releasing workspace");
    }
};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int scratch_address(void **out)
{
    int dev = 0;
    GT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(GT_ERR_INVALID, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        GT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) return fail(GT_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return GT_OK;
}

struct DeviceBuffer {
    void *ptr = nullptr;
    ~DeviceBuffer()
    {
        if (ptr) (void)hipFree(ptr);
    }
    int alloc(size_t bytes, const char *what)
    {
        if (bytes == 0) return GT_OK;
        hipError_t e = device_malloc(&ptr, bytes);
        if (e != hipSuccess) {
            ptr = nullptr;
            return hip_fail(e, what);
        }
        return GT_OK;
    }
};

int dist_method_from_name(const char *name)
{
    // getDistEnum (src/rinterface.cpp:147-156): unknown names -> euclidean.
    if (name == nullptr) return GT_EUCLIDEAN;
    if (strcmp(name, "maximum") == 0) return GT_MAXIMUM;
    if (strcmp(name, "manhattan") == 0) return GT_MANHATTAN;
    if (strcmp(name, "canberra") == 0) return GT_CANBERRA;
    if (strcmp(name, "binary") == 0) return GT_BINARY;
    if (strcmp(name, "minkowski") == 0) return GT_MINKOWSKI;
    return GT_EUCLIDEAN;
}

int dist_host(const float *x_host, int64_t n, int64_t k, int method, float p, int precision,
              int layout, float *out_host)
{
    const int64_t ldd = n;
    int s = check_dist_args(x_host, n, k, k, method, p, precision, layout, out_host, ldd);
    if (s != GT_OK) return s;
    const size_t out_count =
        (layout == GT_LAYOUT_PACKED) ? (size_t)(n > 1 ? n * (n - 1) / 2 : 0) : (size_t)(n * n);
    if (out_count == 0) return GT_OK;
    size_t ws_bytes = 0;
    workspace_bytes(n, k, method, precision, &ws_bytes);
    DeviceBuffer dx, dout, dws;
    if ((s = dx.alloc((size_t)n * (size_t)k * sizeof(float), "hipMalloc(points)")) != GT_OK) return s;
    if ((s = dout.alloc(out_count * sizeof(float), "hipMalloc(distances)")) != GT_OK) return s;
    if ((s = dws.alloc(ws_bytes, "hipMalloc(workspace)")) != GT_OK) return s;
    if (k > 0)
        GT_HIP(hipMemcpy(dx.ptr, x_host, (size_t)n * (size_t)k * sizeof(float),
                         hipMemcpyHostToDevice));
    s = run_dist(static_cast<const float *>(dx.ptr), n, k, k, method, p, precision, layout,
                 static_cast<float *>(dout.ptr), ldd, dws.ptr, nullptr);
    if (s != GT_OK) return s;
    GT_HIP(hipMemcpy(out_host, dout.ptr, out_count * sizeof(float), hipMemcpyDeviceToHost));
    return GT_OK;
}

// ------------------------------ clustering ---------------------------------

int hclust_method_from_name(const char *name)
{
    // getClusterEnum (src/rinterface.cpp:158-170): exact names, anything else
    // (including R's "flexible group", which gpuHclust passes) -> single.
    if (name == nullptr) return GT_HC_SINGLE;
    if (strcmp(name, "complete") == 0) return GT_HC_COMPLETE;
    if (strcmp(name, "wpgma") == 0) return GT_HC_WPGMA;
    if (strcmp(name, "average") == 0) return GT_HC_AVERAGE;
    if (strcmp(name, "median") == 0) return GT_HC_MEDIAN;
    if (strcmp(name, "centroid") == 0) return GT_HC_CENTROID;
    if (strcmp(name, "flexible_group") == 0) return GT_HC_FLEXIBLE_GROUP;
    if (strcmp(name, "flexible") == 0) return GT_HC_FLEXIBLE;
    if (strcmp(name, "ward") == 0) return GT_HC_WARD;
    if (strcmp(name, "mcquitty") == 0) return GT_HC_MCQUITTY;
    return GT_HC_SINGLE;
}

size_t hclust_ws_bytes(int64_t n)
{
    return 3 * round_up((size_t)(n > 0 ? n : 0) * 4, 256);
}

int check_hclust_args(int64_t n, int64_t ldd, int method)
{
    if (n < 0) return fail(GT_ERR_INVALID, "negative size n=%lld", (long long)n);
    if (n >= (int64_t(1) << 24))
        return fail(GT_ERR_INVALID, "clustering supports n below 2^24 (n=%lld)", (long long)n);
    if (ldd < n) return fail(GT_ERR_INVALID, "ldd=%lld < n=%lld", (long long)ldd, (long long)n);
    if (method < GT_HC_SINGLE || method > GT_HC_MCQUITTY)
        return fail(GT_ERR_INVALID, "unknown clustering method %d", method);
    return GT_OK;
}

int run_hclust(float *d, int64_t n, int64_t ldd, int method, float lambda, float beta, int *sub,
               int *sup, float *height, void *ws, hipStream_t stream,
               int state = gt::HC_STATE_AUTO)
{
    char *w = static_cast<char *>(ws);
    const size_t part = round_up((size_t)n * 4, 256);
    GT_HIP(gt::launch_hclust(d, n, ldd, method, lambda, beta, reinterpret_cast<float *>(w),
                             reinterpret_cast<float *>(w + part), reinterpret_cast<int *>(w + 2 * part),
                             sub, sup, height, state, stream));
    return GT_OK;
}

// Merges and heights from a device-resident full matrix (consumed) to host.
int hclust_device_to_host(float *d, int64_t n, int64_t ldd, int method, float lambda, float beta,
                          int *sub_h, int *sup_h, float *height_h)
{
    if (n < 2) return GT_OK;
    const size_t len = (size_t)(n - 1);
    DeviceBuffer ws, out;
    int s;
    if ((s = ws.alloc(hclust_ws_bytes(n), "hipMalloc(clustering workspace)")) != GT_OK) return s;
    if ((s = out.alloc(3 * round_up(len * 4, 256), "hipMalloc(clustering result)")) != GT_OK) return s;
    char *o = static_cast<char *>(out.ptr);
    const size_t part = round_up(len * 4, 256);
    int *dsub = reinterpret_cast<int *>(o);
    int *dsup = reinterpret_cast<int *>(o + part);
    float *dh = reinterpret_cast<float *>(o + 2 * part);
    if ((s = run_hclust(d, n, ldd, method, lambda, beta, dsub, dsup, dh, ws.ptr, nullptr)) != GT_OK)
        return s;
    GT_HIP(hipMemcpy(sub_h, dsub, len * 4, hipMemcpyDeviceToHost));
    GT_HIP(hipMemcpy(sup_h, dsup, len * 4, hipMemcpyDeviceToHost));
    GT_HIP(hipMemcpy(height_h, dh, len * 4, hipMemcpyDeviceToHost));
    return GT_OK;
}

// R's hclust merge matrix and order from the merge sequence, as
// formatClustering / getPrintOrder / depthFirst (src/rinterface.cpp:241-292)
// define them.  The reference rewrites every later reference to a merged
// point in an O(len^2) sweep; the same result in O(len): a point that
// survived step i stands for cluster i + 1 from then on.  Rows list a
// singleton before a cluster and two clusters in increasing order.  The
// order is the leaves of a depth-first walk from the last merge that emits
// the right subtree's leaves from the end of the array backwards first.
int hclust_format(int64_t len, const int *sub, const int *sup, int *merge, int *order)
{
    if (len < 0) return fail(GT_ERR_INVALID, "negative length");
    if (len == 0) {
        if (order) order[0] = 1;
        return GT_OK;
    }
    const int64_t n = len + 1;
    std::vector<int> id((size_t)n, 0);
    for (int64_t i = 0; i < len; ++i) {
        const int p = sub[i], q = sup[i];
        if (p < 0 || p >= n || q < 0 || q >= n || p == q)
            return fail(GT_ERR_INVALID, "invalid merge %lld: (%d, %d)", (long long)i, p, q);
        int a = id[p] ? id[p] : -(p + 1);
        int b = id[q] ? id[q] : -(q + 1);
        if ((b < 0 && a > 0) || (a > 0 && b > 0 && a > b)) std::swap(a, b);
        merge[i] = a;
        merge[i + len] = b;
        id[q] = (int)(i + 1);
        id[p] = (int)(i + 1);
    }
    std::vector<int> stack;
    stack.reserve((size_t)n);
    stack.push_back((int)len);   // cluster ids are 1-based
    int64_t otop = len;
    while (!stack.empty()) {
        const int node = stack.back();
        stack.pop_back();
        if (node < 0) {
            if (otop < 0) return fail(GT_ERR_INVALID, "merge sequence is not a tree");
            order[otop--] = -node;
            continue;
        }
        // right subtree first (it fills the end of the array), then left
        stack.push_back(merge[node - 1]);
        stack.push_back(merge[node - 1 + len]);
    }
    return GT_OK;
}

// ------------------------------ correlation --------------------------------

int check_pmcc_args(int use, int precision, const float *a, int64_t na, int64_t lda,
                    const float *b, int64_t nb, int64_t ldb, int64_t dim, const float *pairs,
                    const float *corr, const float *ts, int64_t ldo)
{
    if (use != GT_USE_EVERYTHING && use != GT_USE_PAIRWISE)
        return fail(GT_ERR_INVALID, "unknown use method %d", use);
    if (precision < GT_PREC_FP32 || precision > GT_PREC_FP16X3)
        return fail(GT_ERR_INVALID, "unknown precision %d", precision);
    if (use == GT_USE_PAIRWISE && precision != GT_PREC_FP32)
        return fail(GT_ERR_UNSUPPORTED,
                    "pairwise.complete.obs runs in FP32 (per-pair means are not a matrix product)");
    if (na < 0 || nb < 0 || dim < 0)
        return fail(GT_ERR_INVALID, "negative size (na=%lld, nb=%lld, dim=%lld)", (long long)na,
                    (long long)nb, (long long)dim);
    // sample counts are carried as FP32 values (the reference's numPairs)
    if (dim >= (int64_t(1) << 24))
        return fail(GT_ERR_INVALID, "dim=%lld: sample counts must stay below 2^24", (long long)dim);
    if (na > (int64_t)std::numeric_limits<int>::max() || nb > (int64_t)std::numeric_limits<int>::max())
        return fail(GT_ERR_INVALID, "too many variables");
    if (lda < dim || (b != nullptr && ldb < dim))
        return fail(GT_ERR_INVALID, "leading dimension below dim");
    if (ldo < nb) return fail(GT_ERR_INVALID, "ldo=%lld < nb=%lld", (long long)ldo, (long long)nb);
    if (na > 0 && nb > 0) {
        if (!pairs || !corr || !ts) return fail(GT_ERR_INVALID, "null output pointer");
        if (dim > 0 && !a) return fail(GT_ERR_INVALID, "null input pointer");
    }
    return GT_OK;
}

// b == nullptr, or b == a with the same shape: correlations of a with itself
// (symmetric result, half the tiles).
bool pmcc_same(const float *a, int64_t na, int64_t lda, const float *b, int64_t nb, int64_t ldb)
{
    return b == nullptr || (b == a && nb == na && ldb == lda);
}

int pmcc_host(int use, int precision, const float *a_host, int64_t na, const float *b_host,
              int64_t nb, int64_t dim, float *pairs, float *corr, float *ts)
{
    int s = check_pmcc_args(use, precision, a_host, na, dim, b_host, nb, dim, dim, pairs, corr,
                            ts, nb);
    if (s != GT_OK) return s;
    if (na == 0 || nb == 0) return GT_OK;
    // gpuCor(x) passes x twice (R/gpuCor.R:7-9): equal contents are the
    // symmetric case, as equal pointers are for pmcc (src/correlation.cpp:73)
    const bool same = b_host == nullptr || b_host == a_host ||
                      (na == nb && memcmp(a_host, b_host, (size_t)na * (size_t)dim * sizeof(float)) == 0);
    const size_t out = (size_t)na * (size_t)nb;
    const size_t ws_bytes = gt::pmcc_workspace_bytes(use, precision, na, nb, same, dim);
    DeviceBuffer da, db, dout, dws;
    const size_t abytes = (size_t)na * (size_t)dim * sizeof(float);
    const size_t bbytes = (size_t)nb * (size_t)dim * sizeof(float);
    if ((s = da.alloc(abytes, "hipMalloc(variables a)")) != GT_OK) return s;
    if (!same && (s = db.alloc(bbytes, "hipMalloc(variables b)")) != GT_OK) return s;
    if ((s = dout.alloc(3 * round_up(out * sizeof(float), 256), "hipMalloc(correlations)")) != GT_OK)
        return s;
    if ((s = dws.alloc(ws_bytes, "hipMalloc(workspace)")) != GT_OK) return s;
    if (abytes) GT_HIP(hipMemcpy(da.ptr, a_host, abytes, hipMemcpyHostToDevice));
    if (!same && bbytes) GT_HIP(hipMemcpy(db.ptr, b_host, bbytes, hipMemcpyHostToDevice));
    char *o = static_cast<char *>(dout.ptr);
    const size_t part = round_up(out * sizeof(float), 256);
    float *dp = reinterpret_cast<float *>(o), *dc = reinterpret_cast<float *>(o + part),
          *dt = reinterpret_cast<float *>(o + 2 * part);
    const float *dbp = same ? static_cast<const float *>(da.ptr) : static_cast<const float *>(db.ptr);
    GT_HIP(gt::launch_pmcc(use, precision, static_cast<const float *>(da.ptr), na, dim, dbp, nb, dim,
                           same, dim, dp, dc, dt, nb, dws.ptr, nullptr));
    GT_HIP(hipMemcpy(pairs, dp, out * sizeof(float), hipMemcpyDeviceToHost));
    GT_HIP(hipMemcpy(corr, dc, out * sizeof(float), hipMemcpyDeviceToHost));
    GT_HIP(hipMemcpy(ts, dt, out * sizeof(float), hipMemcpyDeviceToHost));
    return GT_OK;
}

int ttest_host(const float *pairs, const float *corr, int64_t n, float *ts)
{
    if (n < 0) return fail(GT_ERR_INVALID, "negative size");
    if (n == 0) return GT_OK;
    if (!pairs || !corr || !ts) return fail(GT_ERR_INVALID, "null pointer");
    DeviceBuffer d;
    int s;
    const size_t part = round_up((size_t)n * sizeof(float), 256);
    if ((s = d.alloc(3 * part, "hipMalloc(t-test)")) != GT_OK) return s;
    char *o = static_cast<char *>(d.ptr);
    float *dp = reinterpret_cast<float *>(o), *dc = reinterpret_cast<float *>(o + part),
          *dt = reinterpret_cast<float *>(o + 2 * part);
    GT_HIP(hipMemcpy(dp, pairs, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    GT_HIP(hipMemcpy(dc, corr, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    GT_HIP(gt::launch_ttest(dp, dc, dt, n, nullptr));
    GT_HIP(hipMemcpy(ts, dt, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    return GT_OK;
}

int check_kendall_args(const float *a, int64_t na, int64_t lda, const float *b, int64_t nb,
                       int64_t ldb, int64_t n, const double *tau, int64_t ldt)
{
    if (na < 0 || nb < 0 || n < 0)
        return fail(GT_ERR_INVALID, "negative size (na=%lld, nb=%lld, n=%lld)", (long long)na,
                    (long long)nb, (long long)n);
    // sample-pair words are addressed with 64-bit offsets; 2^22 samples give
    // 2^37 sample pairs per variable
    if (n > (int64_t(1) << 22))
        return fail(GT_ERR_INVALID, "n=%lld samples exceeds the supported 2^22", (long long)n);
    if (lda < n || (b != nullptr && ldb < n)) return fail(GT_ERR_INVALID, "leading dimension below n");
    if (ldt < na) return fail(GT_ERR_INVALID, "ldt=%lld < na=%lld", (long long)ldt, (long long)na);
    if (na > 0 && nb > 0) {
        if (!tau) return fail(GT_ERR_INVALID, "null output pointer");
        if (n > 0 && !a) return fail(GT_ERR_INVALID, "null input pointer");
    }
    return GT_OK;
}

int kendall_host(const float *a_host, int64_t na, const float *b_host, int64_t nb, int64_t n,
                 double *tau)
{
    int s = check_kendall_args(a_host, na, n, b_host, nb, n, n, tau, na);
    if (s != GT_OK) return s;
    if (na == 0 || nb == 0) return GT_OK;
    // gpuCor passes y = x as a second copy (R/gpuCor.R:7-9, 52-55): equal
    // contents are the symmetric case
    const bool same = b_host == nullptr || b_host == a_host ||
                      (na == nb && memcmp(a_host, b_host, (size_t)na * (size_t)n * sizeof(float)) == 0);
    const size_t abytes = (size_t)na * (size_t)n * sizeof(float);
    const size_t bbytes = (size_t)nb * (size_t)n * sizeof(float);
    const size_t out = (size_t)na * (size_t)nb * sizeof(double);
    DeviceBuffer da, db, dout, dws;
    if ((s = da.alloc(abytes, "hipMalloc(variables a)")) != GT_OK) return s;
    if (!same && (s = db.alloc(bbytes, "hipMalloc(variables b)")) != GT_OK) return s;
    if ((s = dout.alloc(out, "hipMalloc(tau)")) != GT_OK) return s;
    if ((s = dws.alloc(gt::kendall_workspace_bytes(na, nb, same, n), "hipMalloc(workspace)")) != GT_OK)
        return s;
    if (abytes) GT_HIP(hipMemcpy(da.ptr, a_host, abytes, hipMemcpyHostToDevice));
    if (!same && bbytes) GT_HIP(hipMemcpy(db.ptr, b_host, bbytes, hipMemcpyHostToDevice));
    const float *dbp = same ? static_cast<const float *>(da.ptr) : static_cast<const float *>(db.ptr);
    GT_HIP(gt::launch_kendall(static_cast<const float *>(da.ptr), na, n, dbp, nb, n, same, n,
                              static_cast<double *>(dout.ptr), na, dws.ptr, nullptr));
    GT_HIP(hipMemcpy(tau, dout.ptr, out, hipMemcpyDeviceToHost));
    return GT_OK;
}

}  // namespace

extern "C" {

const char *gt_last_error(void) { return g_err; }

int gt_version(void) { return 10000; }

int gt_method_from_name(const char *name) { return dist_method_from_name(name); }

int gt_dist_workspace_size(int64_t n, int64_t k, int method, int precision, size_t *bytes)
{
    if (bytes == nullptr) return fail(GT_ERR_INVALID, "null bytes pointer");
    if (n < 0 || k < 0) return fail(GT_ERR_INVALID, "negative size");
    if (method < GT_EUCLIDEAN || method > GT_MINKOWSKI)
        return fail(GT_ERR_INVALID, "unknown distance method %d", method);
    return workspace_bytes(n, k, method, precision, bytes);
}

int gt_dist_ws(const float *x, int64_t n, int64_t k, int64_t ldx, int method, float p,
               int precision, int layout, float *out, int64_t ldd, void *workspace,
               size_t workspace_bytes_given, void *stream)
{
    int s = check_dist_args(x, n, k, ldx, method, p, precision, layout, out, ldd);
    if (s != GT_OK) return s;
    size_t need = 0;
    workspace_bytes(n, k, method, precision, &need);
    if (need > 0 && (workspace == nullptr || workspace_bytes_given < need))
        return fail(GT_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need,
                    workspace_bytes_given);
    return run_dist(x, n, k, ldx, method, p, precision, layout, out, ldd, workspace,
                    static_cast<hipStream_t>(stream));
}

int gt_dist(const float *x, int64_t n, int64_t k, int64_t ldx, int method, float p,
            int precision, int layout, float *out, int64_t ldd, void *stream)
{
    int s = check_dist_args(x, n, k, ldx, method, p, precision, layout, out, ldd);
    if (s != GT_OK) return s;
    size_t need = 0;
    workspace_bytes(n, k, method, precision, &need);
    hipStream_t hs = static_cast<hipStream_t>(stream);
    Scratch ws;
    if ((s = ws.acquire(need, hs)) != GT_OK) return s;
    s = run_dist(x, n, k, ldx, method, p, precision, layout, out, ldd, ws.ptr, hs);
    const int r = ws.release();
    return s != GT_OK ? s : r;
}

int gt_dist_noalloc(const float *x, int64_t n, int64_t k, int64_t ldx, int method, float p,
                    int precision, int layout, float *out, int64_t ldd, void *stream)
{
    int s = check_dist_args(x, n, k, ldx, method, p, precision, layout, out, ldd);
    if (s != GT_OK) return s;
    size_t need = 0;
    workspace_bytes(n, k, method, precision, &need);
    if (need > GT_NOALLOC_SCRATCH_BYTES)
        return fail(GT_ERR_WORKSPACE,
                    "gt_dist_noalloc: this call needs %zu bytes of workspace, above the %zu of "
                    "its static scratch; call gt_dist_ws with a workspace instead",
                    need, (size_t)GT_NOALLOC_SCRATCH_BYTES);
    void *ws = nullptr;
    if (need > 0 && (s = scratch_address(&ws)) != GT_OK) return s;
    // every kernel on this path writes its workspace in full before reading
    // it (split terms and padding, norms, bit rows and counts), so the
    // previous call's contents need no clearing
    return run_dist(x, n, k, ldx, method, p, precision, layout, out, ldd, ws,
                    static_cast<hipStream_t>(stream));
}

int gt_dist_cross(const float *a, int64_t na, int64_t lda, const float *b, int64_t nb,
                  int64_t ldb, int64_t k, int method, float p, float *out, int64_t ldd,
                  void *stream)
{
    int s = check_method(method, p);
    if (s != GT_OK) return s;
    if (na < 0 || nb < 0 || k < 0) return fail(GT_ERR_INVALID, "negative size");
    if (lda < k || ldb < k) return fail(GT_ERR_INVALID, "leading dimension < k");
    if (ldd < na) return fail(GT_ERR_INVALID, "ldd < na");
    if (na == 0 || nb == 0) return GT_OK;
    if (out == nullptr || (k > 0 && (a == nullptr || b == nullptr)))
        return fail(GT_ERR_INVALID, "null pointer");
    hipStream_t hs = static_cast<hipStream_t>(stream);
    if (k == 0) {
        for (int64_t y = 0; y < nb; ++y) GT_HIP(gt::launch_fill(out + y * ldd, na, __builtin_nanf(""), hs));
        return GT_OK;
    }
    if (method == GT_BINARY) {
        const int64_t words = gt::binary_words(k), ldw = gt::binary_ldw(k);
        const size_t abytes = round_up((size_t)na * ldw * sizeof(uint32_t), 256);
        const size_t bbytes = round_up((size_t)nb * ldw * sizeof(uint32_t), 256);
        Scratch ws;
        if ((s = ws.acquire(abytes + bbytes, hs)) != GT_OK) return s;
        uint32_t *abits = static_cast<uint32_t *>(ws.ptr);
        uint32_t *bbits = reinterpret_cast<uint32_t *>(static_cast<char *>(ws.ptr) + abytes);
        hipError_t e = gt::launch_binary_pack(a, na, k, lda, abits, ldw, hs);
        if (e == hipSuccess) e = gt::launch_binary_pack(b, nb, k, ldb, bbits, ldw, hs);
        if (e == hipSuccess)
            e = gt::launch_binary(gt::OUT_CROSS, abits, na, bbits, nb, words, ldw, out, ldd, hs);
        const int r = ws.release();
        if (e != hipSuccess) return hip_fail(e, "binary cross launch");
        return r;
    }
    GT_HIP(gt::launch_plain(plain_op(method, p), gt::OUT_CROSS, a, na, lda, b, nb, ldb, k, p, out,
                            ldd, hs));
    return GT_OK;
}

int gt_dist_host(const float *x_host, int64_t n, int64_t k, int method, float p, int precision,
                 int layout, float *out_host)
{
    return dist_host(x_host, n, k, method, p, precision, layout, out_host);
}

int gt_dist_device_full(const float *x_host, int64_t n, int64_t k, int method, float p,
                        float **d_out, size_t *pitch_bytes)
{
    if (d_out == nullptr || pitch_bytes == nullptr) return fail(GT_ERR_INVALID, "null output pointer");
    *d_out = nullptr;
    *pitch_bytes = 0;
    int s = check_dist_args(x_host, n, k, k, method, p, GT_PREC_FP32, GT_LAYOUT_FULL,
                            reinterpret_cast<const float *>(d_out), n);
    if (s != GT_OK) return s;
    if (n == 0) return GT_OK;
    size_t ws_bytes = 0;
    workspace_bytes(n, k, method, GT_PREC_FP32, &ws_bytes);
    DeviceBuffer dx, dws;
    void *d = nullptr;
    size_t pitch = 0;
    if ((s = dx.alloc((size_t)n * (size_t)k * sizeof(float), "hipMalloc(points)")) != GT_OK) return s;
    if ((s = dws.alloc(ws_bytes, "hipMalloc(workspace)")) != GT_OK) return s;
    // Pitched n x n result, as distanceLeaveOnGpu allocates it (cudaMallocPitch).
    hipError_t e = injected_alloc_failure()
                       ? hipErrorOutOfMemory
                       : hipMallocPitch(&d, &pitch, (size_t)n * sizeof(float), (size_t)n);
    if (e != hipSuccess) return hip_fail(e, "hipMallocPitch(distances)");
    if (k > 0) {
        e = hipMemcpy(dx.ptr, x_host, (size_t)n * (size_t)k * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d);
            return hip_fail(e, "hipMemcpy(points)");
        }
    }
    s = run_dist(static_cast<const float *>(dx.ptr), n, k, k, method, p, GT_PREC_FP32,
                 GT_LAYOUT_FULL, static_cast<float *>(d), (int64_t)(pitch / sizeof(float)),
                 dws.ptr, nullptr);
    if (s == GT_OK) {
        e = hipDeviceSynchronize();
        if (e != hipSuccess) s = hip_fail(e, "distance kernel");
    }
    if (s != GT_OK) {
        (void)hipFree(d);
        return s;
    }
    *d_out = static_cast<float *>(d);
    *pitch_bytes = pitch;
    return GT_OK;
}

int gt_free(void *device_ptr)
{
    if (device_ptr == nullptr) return GT_OK;
    GT_HIP(hipFree(device_ptr));
    return GT_OK;
}

int gt_hclust_method_from_name(const char *name) { return hclust_method_from_name(name); }

int gt_hclust_workspace_size(int64_t n, size_t *bytes)
{
    if (bytes == nullptr) return fail(GT_ERR_INVALID, "null bytes pointer");
    if (n < 0) return fail(GT_ERR_INVALID, "negative size");
    *bytes = hclust_ws_bytes(n);
    return GT_OK;
}

int gt_hclust(float *d, int64_t n, int64_t ldd, int method, float lambda, float beta, int *sub,
              int *sup, float *height, void *workspace, size_t workspace_bytes, void *stream)
{
    int s = check_hclust_args(n, ldd, method);
    if (s != GT_OK) return s;
    if (n < 2) return GT_OK;
    if (!d || !sub || !sup || !height) return fail(GT_ERR_INVALID, "null pointer");
    if (workspace == nullptr || workspace_bytes < hclust_ws_bytes(n))
        return fail(GT_ERR_WORKSPACE, "clustering workspace of %zu bytes needed (got %zu)",
                    hclust_ws_bytes(n), workspace_bytes);
    return run_hclust(d, n, ldd, method, lambda, beta, sub, sup, height, workspace,
                      static_cast<hipStream_t>(stream));
}

int gt_hclust_state(float *d, int64_t n, int64_t ldd, int method, float lambda, float beta,
                    int *sub, int *sup, float *height, void *workspace, size_t workspace_bytes,
                    int state, void *stream)
{
    int s = check_hclust_args(n, ldd, method);
    if (s != GT_OK) return s;
    if (state != GT_HC_STATE_AUTO && state != GT_HC_STATE_LDS && state != GT_HC_STATE_GLOBAL)
        return fail(GT_ERR_INVALID, "unknown clustering state placement %d", state);
    if (state == GT_HC_STATE_LDS && n > gt::HC_LDS_MAX_N)
        return fail(GT_ERR_INVALID, "the LDS placement holds at most %lld points (n=%lld)",
                    (long long)gt::HC_LDS_MAX_N, (long long)n);
    if (n < 2) return GT_OK;
    if (!d || !sub || !sup || !height) return fail(GT_ERR_INVALID, "null pointer");
    if (workspace == nullptr || workspace_bytes < hclust_ws_bytes(n))
        return fail(GT_ERR_WORKSPACE, "clustering workspace of %zu bytes needed (got %zu)",
                    hclust_ws_bytes(n), workspace_bytes);
    static_assert((int)GT_HC_STATE_LDS == (int)gt::HC_STATE_LDS &&
                  (int)GT_HC_STATE_GLOBAL == (int)gt::HC_STATE_GLOBAL, "state values");
    return run_hclust(d, n, ldd, method, lambda, beta, sub, sup, height, workspace,
                      static_cast<hipStream_t>(stream), state);
}

int gt_hclust_host(const float *d_host, int64_t n, int64_t ldd, int method, float lambda,
                   float beta, int *sub, int *sup, float *height)
{
    int s = check_hclust_args(n, ldd, method);
    if (s != GT_OK) return s;
    if (n < 2) return GT_OK;
    if (!d_host || !sub || !sup || !height) return fail(GT_ERR_INVALID, "null pointer");
    DeviceBuffer dd;
    if ((s = dd.alloc((size_t)n * (size_t)n * sizeof(float), "hipMalloc(distances)")) != GT_OK)
        return s;
    GT_HIP(hipMemcpy2D(dd.ptr, (size_t)n * sizeof(float), d_host, (size_t)ldd * sizeof(float),
                       (size_t)n * sizeof(float), (size_t)n, hipMemcpyHostToDevice));
    return hclust_device_to_host(static_cast<float *>(dd.ptr), n, n, method, lambda, beta, sub, sup,
                                 height);
}

int gt_hclust_format(int64_t len, const int *sub, const int *sup, int *merge, int *order)
{
    if (len < 0 || (len > 0 && (!sub || !sup || !merge)) || !order)
        return fail(GT_ERR_INVALID, "null or negative argument");
    return hclust_format(len, sub, sup, merge, order);
}

int gt_pmcc_workspace_size(int use, int precision, int64_t na, int64_t nb, int64_t dim, int same,
                           size_t *bytes)
{
    if (bytes == nullptr) return fail(GT_ERR_INVALID, "null bytes pointer");
    if (na < 0 || nb < 0 || dim < 0) return fail(GT_ERR_INVALID, "negative size");
    if (use != GT_USE_EVERYTHING && use != GT_USE_PAIRWISE)
        return fail(GT_ERR_INVALID, "unknown use method %d", use);
    if (precision < GT_PREC_FP32 || precision > GT_PREC_FP16X3)
        return fail(GT_ERR_INVALID, "unknown precision %d", precision);
    *bytes = gt::pmcc_workspace_bytes(use, precision, na, same ? na : nb, same != 0, dim);
    return GT_OK;
}

int gt_pmcc(int use, int precision, const float *a, int64_t na, int64_t lda, const float *b,
            int64_t nb, int64_t ldb, int64_t dim, float *pairs, float *corr, float *ts, int64_t ldo,
            void *workspace, size_t workspace_bytes, void *stream)
{
    const bool same = pmcc_same(a, na, lda, b, nb, ldb);
    if (b == nullptr) {
        b = a;
        nb = na;
        ldb = lda;
    }
    int s = check_pmcc_args(use, precision, a, na, lda, b, nb, ldb, dim, pairs, corr, ts, ldo);
    if (s != GT_OK) return s;
    const size_t need = gt::pmcc_workspace_bytes(use, precision, na, nb, same, dim);
    if (need > 0 && (workspace == nullptr || workspace_bytes < need))
        return fail(GT_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need,
                    workspace_bytes);
    GT_HIP(gt::launch_pmcc(use, precision, a, na, lda, b, nb, ldb, same, dim, pairs, corr, ts, ldo,
                           workspace, static_cast<hipStream_t>(stream)));
    return GT_OK;
}

int gt_pmcc_host(int use, int precision, const float *a_host, int64_t na, const float *b_host,
                 int64_t nb, int64_t dim, float *pairs, float *corr, float *ts)
{
    if (b_host == nullptr) nb = na;
    return pmcc_host(use, precision, a_host, na, b_host, nb, dim, pairs, corr, ts);
}

int gt_ttest(const float *pairs, const float *corr, int64_t n, float *ts, void *stream)
{
    if (n < 0) return fail(GT_ERR_INVALID, "negative size");
    if (n > 0 && (!pairs || !corr || !ts)) return fail(GT_ERR_INVALID, "null pointer");
    GT_HIP(gt::launch_ttest(pairs, corr, ts, n, static_cast<hipStream_t>(stream)));
    return GT_OK;
}

int gt_kendall_workspace_size(int64_t na, int64_t nb, int64_t n, int same, size_t *bytes)
{
    if (bytes == nullptr) return fail(GT_ERR_INVALID, "null bytes pointer");
    if (na < 0 || nb < 0 || n < 0) return fail(GT_ERR_INVALID, "negative size");
    *bytes = gt::kendall_workspace_bytes(na, same ? na : nb, same != 0, n);
    return GT_OK;
}

int gt_kendall(const float *a, int64_t na, int64_t lda, const float *b, int64_t nb, int64_t ldb,
               int64_t n, double *tau, int64_t ldt, void *workspace, size_t workspace_bytes,
               void *stream)
{
    const bool same = b == nullptr || (b == a && nb == na && ldb == lda);
    if (b == nullptr) {
        b = a;
        nb = na;
        ldb = lda;
    }
    int s = check_kendall_args(a, na, lda, b, nb, ldb, n, tau, ldt);
    if (s != GT_OK) return s;
    const size_t need = gt::kendall_workspace_bytes(na, nb, same, n);
    if (need > 0 && (workspace == nullptr || workspace_bytes < need))
        return fail(GT_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need,
                    workspace_bytes);
    GT_HIP(gt::launch_kendall(a, na, lda, b, nb, ldb, same, n, tau, ldt, workspace,
                              static_cast<hipStream_t>(stream)));
    return GT_OK;
}

int gt_kendall_host(const float *a_host, int64_t na, const float *b_host, int64_t nb, int64_t n,
                    double *tau)
{
    if (b_host == nullptr) nb = na;
    return kendall_host(a_host, na, b_host, nb, n, tau);
}

// --------------------------- .C drop-in exports ---------------------------

void Rdistances(const float *points, const int *numPoints, const int *dim, float *distances,
                const char **method, const float *p)
{
    g_err[0] = '\0';
    int s;
    if (!numPoints || !dim || !method || !p) {
        s = fail(GT_ERR_INVALID, "Rdistances: null argument");
    } else {
        const int m = dist_method_from_name(*method);
        s = dist_host(points, *numPoints, *dim, m, *p, GT_PREC_FP32, GT_LAYOUT_FULL, distances);
    }
    raise_in_r(s);
}

void Rdistances_packed(const float *points, const int *numPoints, const int *dim,
                       float *distances, const char **method, const float *p)
{
    g_err[0] = '\0';
    int s;
    if (!numPoints || !dim || !method || !p) {
        s = fail(GT_ERR_INVALID, "Rdistances_packed: null argument");
    } else {
        const int m = dist_method_from_name(*method);
        s = dist_host(points, *numPoints, *dim, m, *p, GT_PREC_FP32, GT_LAYOUT_PACKED, distances);
    }
    raise_in_r(s);
}

void Rdistances_packed_prec(const float *points, const int *numPoints, const int *dim,
                            float *distances, const char **method, const float *p,
                            const char **precision)
{
    g_err[0] = '\0';
    int s;
    if (!numPoints || !dim || !method || !p || !precision || !*precision) {
        s = fail(GT_ERR_INVALID, "Rdistances_packed_prec: null argument");
    } else {
        int prec = -1;
        if (strcmp(*precision, "fp32") == 0) prec = GT_PREC_FP32;
        else if (strcmp(*precision, "bf16x3") == 0) prec = GT_PREC_BF16X3;
        else if (strcmp(*precision, "fp16x3") == 0) prec = GT_PREC_FP16X3;
        if (prec < 0) {
            s = fail(GT_ERR_INVALID, "unknown precision \"%s\" (fp32, bf16x3 or fp16x3)", *precision);
        } else {
            const int m = dist_method_from_name(*method);
            s = dist_host(points, *numPoints, *dim, m, *p, prec, GT_LAYOUT_PACKED, distances);
        }
    }
    raise_in_r(s);
}

void Rdistclust(const char **distmethod, const char **clustmethod, const float *points,
                const int *numPoints, const int *dim, int *merge, int *order, float *val)
{
    g_err[0] = '\0';
    int s;
    if (!distmethod || !clustmethod || !numPoints || !dim) {
        s = fail(GT_ERR_INVALID, "Rdistclust: null argument");
    } else if (*numPoints < 1) {
        s = fail(GT_ERR_INVALID, "Rdistclust: need at least one point");
    } else {
        // Distance stage as Rdistclust runs it (src/rinterface.cpp:182-183):
        // a device-resident full matrix, p fixed at 2; then the clustering
        // stage on that matrix and R's merge / order formatting.
        const int64_t n = *numPoints;
        float *d = nullptr;
        size_t pitch = 0;
        s = gt_dist_device_full(points, n, *dim, dist_method_from_name(*distmethod), 2.0f, &d, &pitch);
        if (s == GT_OK) {
            std::vector<int> sub((size_t)(n > 1 ? n - 1 : 0)), sup(sub.size());
            s = check_hclust_args(n, (int64_t)(pitch / sizeof(float)), hclust_method_from_name(*clustmethod));
            if (s == GT_OK)
                s = hclust_device_to_host(d, n, (int64_t)(pitch / sizeof(float)),
                                          hclust_method_from_name(*clustmethod), 0.5f, 0.5f, sub.data(),
                                          sup.data(), val);
            if (s == GT_OK) s = hclust_format(n - 1, sub.data(), sup.data(), merge, order);
        }
        if (d) (void)hipFree(d);
    }
    raise_in_r(s);
}

void Rhcluster(const float *distMat, const int *numPoints, int *merge, int *order, float *val,
               const char **method)
{
    g_err[0] = '\0';
    int s;
    if (!distMat || !numPoints || !merge || !order || !method) {
        s = fail(GT_ERR_INVALID, "Rhcluster: null argument");
    } else if (*numPoints < 1) {
        s = fail(GT_ERR_INVALID, "Rhcluster: need at least one point");
    } else {
        const int64_t n = *numPoints;
        std::vector<int> sub((size_t)(n > 1 ? n - 1 : 0)), sup(sub.size());
        // lambda = beta = 0.5, as Rhcluster passes them (src/rinterface.cpp:226-227)
        s = gt_hclust_host(distMat, n, n, hclust_method_from_name(*method), 0.5f, 0.5f, sub.data(),
                           sup.data(), val);
        if (s == GT_OK) s = hclust_format(n - 1, sub.data(), sup.data(), merge, order);
    }
    raise_in_r(s);
}

void rpmcc(const int *whichObs, const float *samplesA, const int *numSamplesA,
           const float *samplesB, const int *numSamplesB, const int *sampleSize, float *numPairs,
           float *correlations, float *signifs)
{
    g_err[0] = '\0';
    int s;
    if (!whichObs || !numSamplesA || !numSamplesB || !sampleSize) {
        s = fail(GT_ERR_INVALID, "rpmcc: null argument");
    } else if (*whichObs != 0 && *whichObs != 1) {
        s = fail(GT_ERR_INVALID, "unknown use method");   // rinterface.cpp:47-48
    } else {
        // "everything" on the matrix cores (BF16 three-term split: products
        // within 2^-23 of FP32's), "pairwise.complete.obs" in FP32
        const int use = *whichObs;
        const int prec = use == GT_USE_EVERYTHING ? GT_PREC_BF16X3 : GT_PREC_FP32;
        s = pmcc_host(use, prec, samplesA, *numSamplesA, samplesB, *numSamplesB, *sampleSize,
                      numPairs, correlations, signifs);
    }
    raise_in_r(s);
}

void rtestT(const float *pairs, const float *coeffs, const int *n, float *ts)
{
    g_err[0] = '\0';
    const int s = n ? ttest_host(pairs, coeffs, *n, ts) : fail(GT_ERR_INVALID, "rtestT: null argument");
    raise_in_r(s);
}

void RgpuKendall(const float *X, const int *nx, const float *Y, const int *ny,
                 const int *sampleSize, double *answers)
{
    g_err[0] = '\0';
    int s;
    if (!nx || !ny || !sampleSize)
        s = fail(GT_ERR_INVALID, "RgpuKendall: null argument");
    else
        s = kendall_host(X, *nx, Y, *ny, *sampleSize, answers);
    raise_in_r(s);
}

void cuCompile(const int *numFiles, const char **cuFilenames, const char **cuSrc)
{
    (void)numFiles;
    (void)cuFilenames;
    (void)cuSrc;
    package_loaded = true;
}

void unloadPackage(void) { package_loaded = false; }

void rsetDevice(const int *device)
{
    g_err[0] = '\0';
    int s = GT_OK;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) {
        s = hip_fail(e, "hipGetDeviceCount");
    } else if (device == nullptr || *device < 0 || *device >= count) {
        s = fail(GT_ERR_INVALID, "The gpu id number is not valid.");
    } else {
        e = hipSetDevice(*device);
        if (e != hipSuccess) s = hip_fail(e, "setDevice function : choosing gpu");
    }
    raise_in_r(s);
}

void rgetDevice(int *device)
{
    g_err[0] = '\0';
    int s = GT_OK;
    if (device == nullptr) {
        s = fail(GT_ERR_INVALID, "rgetDevice: null argument");
    } else {
        hipError_t e = hipGetDevice(device);
        if (e != hipSuccess) s = hip_fail(e, "hipGetDevice");
    }
    raise_in_r(s);
}

}  // extern "C"

// ------------------------------ gradients ---------------------------------
extern "C" {
int gt_dist_backward(const float *x, int64_t n, int64_t k, int64_t ldx, int method, float p,
                     const float *g_packed, float *dx, int64_t lddx, void *stream)
{
    if (method != GT_EUCLIDEAN && method != GT_MINKOWSKI)
        return fail(GT_ERR_UNSUPPORTED, "backward exists for euclidean and minkowski (method %d)",
                    method);
    if (method == GT_MINKOWSKI && !(p >= 1.0f && isfinite(p)))
        return fail(GT_ERR_INVALID, "minkowski backward needs a finite p >= 1 (p = %g)", (double)p);
    if (n < 0 || k < 0) return fail(GT_ERR_INVALID, "negative size");
    if (ldx < k || lddx < k) return fail(GT_ERR_INVALID, "leading dimension < k");
    if (n == 0 || k == 0) return GT_OK;
    if (x == nullptr || dx == nullptr || (n > 1 && g_packed == nullptr))
        return fail(GT_ERR_INVALID, "null pointer");
    hipStream_t hs = static_cast<hipStream_t>(stream);
    if (n == 1) {
        GT_HIP(gt::launch_fill(dx, k, 0.0f, hs));
        return GT_OK;
    }
    int op = gt::OP_MINKP;
    if (method == GT_EUCLIDEAN || p == 2.0f) op = gt::OP_EUCLID;
    else if (p == 3.0f) op = gt::OP_MINK3;
    // partial dX slices when the row blocks are few (stream-ordered scratch)
    Scratch ws;
    int s = ws.acquire(gt::backward_workspace_bytes(n, k), hs);
    if (s != GT_OK) return s;
    const hipError_t e = gt::launch_backward(op, x, n, k, ldx, p, g_packed, dx, lddx,
                                             static_cast<float *>(ws.ptr), hs);
    const int r = ws.release();
    if (e != hipSuccess) return hip_fail(e, "gt_dist_backward");
    return r;
}

int gt_mds_stress(const float *y, int64_t n, int64_t q, int64_t ldy, const float *delta_packed,
                  float *dy, int64_t lddy, double *stress_dev, void *stream)
{
    if (n < 0 || q < 1) return fail(GT_ERR_INVALID, "need n >= 0 and q >= 1");
    if (ldy < q || (dy != nullptr && lddy < q)) return fail(GT_ERR_INVALID, "leading dimension < q");
    if (stress_dev == nullptr) return fail(GT_ERR_INVALID, "null stress pointer");
    if (n > 1 && (y == nullptr || delta_packed == nullptr)) return fail(GT_ERR_INVALID, "null pointer");
    hipStream_t hs = static_cast<hipStream_t>(stream);
    const int64_t parts = n > 1 ? gt::mds_partials_count(n, q) : 1;
    const size_t pbytes = round_up((size_t)parts * sizeof(double), 256);
    const size_t dbytes = (n > 1 && dy != nullptr) ? gt::backward_workspace_bytes(n, q) : 0;
    Scratch ws;
    int s = ws.acquire(pbytes + dbytes, hs);
    if (s != GT_OK) return s;
    hipError_t e;
    if (n < 2) {
        // a fill kernel (not a memset node) so that the call stays capturable
        e = gt::launch_fill(reinterpret_cast<float *>(stress_dev), 2, 0.0f, hs);
        if (e == hipSuccess && dy != nullptr && n == 1) e = gt::launch_fill(dy, q, 0.0f, hs);
    } else {
        e = gt::launch_mds(y, n, q, ldy, delta_packed, dy, lddy, stress_dev,
                           static_cast<double *>(ws.ptr),
                           dbytes ? reinterpret_cast<float *>(static_cast<char *>(ws.ptr) + pbytes)
                                  : nullptr,
                           hs);
    }
    const int r = ws.release();
    if (e != hipSuccess) return hip_fail(e, "gt_mds_stress");
    return r;
}

int gt_batched_tile_count(void) { return 3; }

int gt_dist_batched_schedule(const int64_t *sizes_host, const int64_t *dims_host, int64_t nsets,
                             int tile, int64_t *schedule_host, int64_t capacity, int64_t *count)
{
    if (count == nullptr) return fail(GT_ERR_INVALID, "null count pointer");
    *count = 0;
    const int bm = gt::batched_tile_side(tile);
    if (bm == 0) return fail(GT_ERR_INVALID, "unknown tile shape %d", tile);
    if (nsets < 0 || nsets > (int64_t)INT32_MAX) return fail(GT_ERR_INVALID, "bad number of sets");
    if (nsets > 0 && (sizes_host == nullptr || dims_host == nullptr))
        return fail(GT_ERR_INVALID, "null descriptor pointer");
    // entries: (set, I, J) for every lower tile J <= I of every set with >= 2
    // rows and >= 1 feature; cost order: more features first, then bigger sets
    struct Ent { int64_t key, d, n; };
    std::vector<Ent> ents;
    for (int64_t s = 0; s < nsets; ++s) {
        const int64_t n = sizes_host[s], d = dims_host[s];
        if (n < 0 || d < 0) return fail(GT_ERR_INVALID, "set %lld has a negative size", (long long)s);
        if (n < 2) continue;
        if (d < 1) return fail(GT_ERR_INVALID, "set %lld has no features", (long long)s);
        const int64_t T = (n + bm - 1) / bm;
        if (T > 65535) return fail(GT_ERR_INVALID, "set %lld is too large for batching", (long long)s);
        for (int64_t I = 0; I < T; ++I)
            for (int64_t J = 0; J <= I; ++J) ents.push_back({(s << 32) | (I << 16) | J, d, n});
    }
    std::stable_sort(ents.begin(), ents.end(), [](const Ent &a, const Ent &b) {
        return a.d != b.d ? a.d > b.d : a.n > b.n;
    });
    *count = (int64_t)ents.size();
    if (schedule_host != nullptr && capacity >= *count)
        for (size_t e = 0; e < ents.size(); ++e) schedule_host[e] = ents[e].key;
    return GT_OK;
}

int gt_dist_batched(const float *x, const int64_t *x_offsets, const int64_t *sizes,
                    const int64_t *dims, int64_t nsets, float *out, const int64_t *out_offsets,
                    int method, int tile, const int64_t *schedule, int64_t count, void *stream)
{
    int op;
    switch (method) {
    case GT_EUCLIDEAN: op = gt::OP_EUCLID; break;
    case GT_MANHATTAN: op = gt::OP_MANHATTAN; break;
    case GT_MAXIMUM: op = gt::OP_MAXIMUM; break;
    default:
        return fail(GT_ERR_UNSUPPORTED, "batched path exists for euclidean, manhattan, maximum");
    }
    if (gt::batched_tile_side(tile) == 0) return fail(GT_ERR_INVALID, "unknown tile shape %d", tile);
    if (count < 0 || nsets < 0) return fail(GT_ERR_INVALID, "negative count");
    if (count == 0) return GT_OK;
    if (!x || !x_offsets || !sizes || !dims || !out || !out_offsets || !schedule)
        return fail(GT_ERR_INVALID, "null pointer");
    int dev = 0, cus = 0;
    GT_HIP(hipGetDevice(&dev));
    GT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int64_t grid = (int64_t)cus * gt::batched_resident_blocks(op, tile);
    if (grid > count) grid = count;
    if (grid < 1) grid = 1;
    GT_HIP(gt::launch_batched(op, tile, x, x_offsets, sizes, dims, out_offsets, out, schedule, count,
                              nsets, (int)grid, static_cast<hipStream_t>(stream)));
    return GT_OK;
}
}
