// noalloc.hip -- the fit stage for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops): qcs_fit_dev_noalloc.
//
// qcs_fit_dev carves the multi-emitter cascade's scratch (lists, flags, scan offsets,
// second-emitter rows, counters) from a grow-only buffer of the context, which may call
// hipMalloc, and reads the row count back with a pinned copy and a stream synchronisation.  This
// variant takes the same scratch from one file-scope device array of fixed size instead, so a
// call allocates nothing, and leaves the row count on the device.  Each GPU holds its own copy
// of the array (SCRATCH_BYTES) from the first call on it until the process ends; host code uses
// the address hipGetSymbolAddress returns for the current device, never the array's name.
// Calls that overlap on one GPU share the array and race: serialise them (one stream, or one
// caller at a time per device).
#include <mutex>
#include "qcs_internal.h"

namespace {
// cascade scratch at QCS_NOALLOC_MAX_ROIS ROIs is 13,926,912 bytes (cascade_scratch_bytes)
constexpr size_t SCRATCH_BYTES = (size_t)14 << 20;
constexpr int MAX_DEVICES = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace qcs {

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[MAX_DEVICES];

// Resolve this device's copy of `scratch` (the device must be current); once per device.
int noalloc_scratch_init(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return fail(QCS_ERR_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return QCS_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    if (!ptr) return fail(QCS_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[device] = ptr;
    return QCS_OK;
}

static void *scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

__global__ void copy_count_kernel(const int *__restrict__ src, int *__restrict__ dst) { *dst = *src; }

}  // namespace qcs

using namespace qcs;

extern "C" int qcs_fit_dev_noalloc(qcs_ctx *c, const uint16_t *d_rois, const float *d_wle, int n, float *d_loc, int loc_cap,
                                   int *d_n_rows, void *stream)
{
    if (!c || n < 0 || (n > 0 && (!d_rois || !d_loc))) return fail(QCS_ERR_ARG, "invalid argument");
    if (n > QCS_NOALLOC_MAX_ROIS)
        return fail(QCS_ERR_CAPACITY, "%d ROIs exceed the allocation-free limit of %d per call", n, QCS_NOALLOC_MAX_ROIS);
    if (n > 0 && loc_cap < n) return fail(QCS_ERR_CAPACITY, "loc_cap (%d) < number of ROIs (%d)", loc_cap, n);
    if (n > 0 && !d_wle) return fail(QCS_ERR_ARG, "WLE records are required (qcs_wle_dev computes them)");
    int dev = -1;
    QCS_HIP(hipGetDevice(&dev));
    if (dev != c->device) return fail(QCS_ERR_STATE, "context belongs to device %d, current device is %d", c->device, dev);
    int rc = noalloc_scratch_init(dev);
    if (rc) return rc;
    void *base = scratch_ptr(dev);
    hipStream_t s = (hipStream_t)stream;
    if (n == 0) {
        if (d_n_rows) fill_async(d_n_rows, 1, 0, s);
        QCS_HIP(hipGetLastError());
        return QCS_OK;
    }
    if (!c->p.MultiEmitterFitEn) {   // one-emitter mode: one row per ROI, no scratch
        if ((rc = fit_launch_1e(c->p, d_rois, d_wle, nullptr, n, nullptr, d_loc, nullptr, s))) return rc;
        if ((rc = crlb_launch(c->p, d_loc, n, nullptr, s))) return rc;
        if (d_n_rows) fill_async(d_n_rows, 1, n, s);
        QCS_HIP(hipGetLastError());
        return QCS_OK;
    }
    if (c->p.ROISize > 13) return fail(QCS_ERR_ARG, "multi-emitter fitting supports ROISize 5..13 (got %d)", c->p.ROISize);
    if (cascade_scratch_bytes(n) > SCRATCH_BYTES) return fail(QCS_ERR_CAPACITY, "cascade scratch exceeds the static array");
    CascadeBufs b;
    cascade_carve(base, n, b);
    if ((rc = fit_cascade_launch(c->p, nullptr, d_rois, d_wle, n, d_loc, loc_cap, b, s))) return rc;
    if (d_n_rows) hipLaunchKernelGGL(copy_count_kernel, dim3(1), dim3(1), 0, s, b.cnt + 7, d_n_rows);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}
