// sim.hip -- deterministic synthetic ROIs and frames (host and gfx950 implementations of
// sim_core.h).  Input generation for tests and benchmarks; not part of the localization path.
#include <math.h>
#include <string.h>
#include <vector>
#include "qcs_internal.h"
#include "sim_core.h"

using namespace qcs;
using qcs_sim::Emitter;

namespace {

__global__ void sim_rois_kernel(qcs_sim_params sp, uint64_t seed, int n, int first_frame, uint16_t *rois, float *truth)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int R = sp.width;
    const int M = sp.n_per_unit_max > 1 ? sp.n_per_unit_max : 1;
    qcs_sim::roi_unit(sp, seed, i, first_frame, rois + (size_t)i * (R * (R + 1)),
                      truth ? truth + (size_t)i * M * QCS_TRUTH_FLOATS : nullptr);
}

// splat emitters of one frame into a double image (atomics; sums are order-independent to 1 ulp)
__global__ void sim_splat_kernel(qcs_sim_params sp, uint64_t seed, int frame, int nem, double *img)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nem) return;
    Emitter e;
    if (!qcs_sim::frame_slot_emitter(sp, seed, frame, k, e)) return;
    const double reach = qcs_sim::emitter_reach(e);
    int c0 = (int)ceil(e.x - reach - 0.5), c1 = (int)floor(e.x + reach - 0.5);
    int r0 = (int)ceil(e.y - reach - 0.5), r1 = (int)floor(e.y + reach - 0.5);
    c0 = c0 < 0 ? 0 : c0;
    r0 = r0 < 0 ? 0 : r0;
    c1 = c1 > sp.width - 1 ? sp.width - 1 : c1;
    r1 = r1 > sp.height - 1 ? sp.height - 1 : r1;
    for (int r = r0; r <= r1; r++)
        for (int c = c0; c <= c1; c++) atomicAdd(&img[(size_t)r * sp.width + c], qcs_sim::psf_value(e, c + 0.5, r + 0.5));
}

__global__ void sim_noise_kernel(qcs_sim_params sp, uint64_t seed, int frame, const double *img, uint16_t *out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t npx = (size_t)sp.width * sp.height;
    if (i >= npx) return;
    const double bg = qcs_sim::frame_background(sp, seed, frame);
    qcs_sim::Gen g(seed, (uint32_t)frame, (uint32_t)i, qcs_sim::TAG_FRAME_PIX);
    const int e = qcs_sim::poisson((double)sp.qe * (bg + img[i]), g);
    out[i] = qcs_sim::to_adu(e, sp);
}

int check_sim(const qcs_sim_params *sp)
{
    if (!sp) return fail(QCS_ERR_ARG, "null simulator parameters");
    if (sp->width <= 0 || sp->height <= 0) return fail(QCS_ERR_ARG, "simulator size must be positive");
    if (sp->n_per_unit_max > 4) return fail(QCS_ERR_ARG, "n_per_unit_max must be <= 4");
    if (sp->density_max_per_px < 0.0f) return fail(QCS_ERR_ARG, "density_max_per_px must be >= 0");
    return QCS_OK;
}

}  // namespace

extern "C" {

void qcs_sim_params_default(qcs_sim_params *sp)
{
    if (!sp) return;
    memset(sp, 0, sizeof(*sp));
    sp->width = 7;
    sp->height = 7;
    sp->offset = 100.0f;
    sp->kadc = 0.45f;
    sp->qe = 1.0f;
    sp->photons_min = 500.0f;
    sp->photons_max = 5000.0f;
    sp->sigma_min = 1.0f;
    sp->sigma_max = 1.6f;
    sp->bg_min = 10.0f;
    sp->bg_max = 100.0f;
    sp->density_per_px = 0.003f;
    sp->center_jitter = 1.0f;
    sp->astig = 0;
    sp->z_min = -500.0f;
    sp->z_max = 500.0f;
    sp->s0 = 1.2f;
    sp->c_nm = 250.0f;
    sp->d_nm = 400.0f;
    sp->n_per_unit_max = 1;
    sp->min_sep = 1.0f;
    sp->max_sep = 3.0f;
    sp->density_max_per_px = 0.0f;
}

int qcs_sim_rois_host(const qcs_sim_params *sp, uint64_t seed, int n, int first_frame, uint16_t *rois, float *truth)
{
    int rc = check_sim(sp);
    if (rc) return rc;
    if (n < 0 || (!rois && n > 0)) return fail(QCS_ERR_ARG, "invalid ROI buffer");
    const int R = sp->width;
    const int M = sp->n_per_unit_max > 1 ? sp->n_per_unit_max : 1;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
        qcs_sim::roi_unit(*sp, seed, i, first_frame, rois + (size_t)i * (R * (R + 1)),
                          truth ? truth + (size_t)i * M * QCS_TRUTH_FLOATS : nullptr);
    return QCS_OK;
}

int qcs_sim_rois_dev(const qcs_sim_params *sp, uint64_t seed, int n, int first_frame, uint16_t *d_rois, float *d_truth,
                     void *stream)
{
    int rc = check_sim(sp);
    if (rc) return rc;
    if (n <= 0) return QCS_OK;
    hipLaunchKernelGGL(sim_rois_kernel, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)stream, *sp, seed, n, first_frame,
                       d_rois, d_truth);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

int qcs_sim_frame_emitters(const qcs_sim_params *sp, uint64_t seed, int frame, float *truth, int cap)
{
    int rc = check_sim(sp);
    if (rc) return -rc;
    const int slots = qcs_sim::frame_slots(*sp, seed, frame);
    const double bg = qcs_sim::frame_background(*sp, seed, frame);
    int n = 0;
    for (int k = 0; k < slots; k++) {
        Emitter e;
        if (!qcs_sim::frame_slot_emitter(*sp, seed, frame, k, e)) continue;
        if (truth && n < cap) {
            float *t = truth + (size_t)n * QCS_TRUTH_FLOATS;
            t[0] = (float)frame; t[1] = (float)e.x; t[2] = (float)e.y; t[3] = (float)e.z;
            t[4] = (float)e.n; t[5] = (float)bg; t[6] = (float)e.sx; t[7] = (float)e.sy;
        }
        n++;
    }
    return n;
}

int qcs_sim_frames_host(const qcs_sim_params *sp, uint64_t seed, int first_frame, int nframes, uint16_t *frames)
{
    int rc = check_sim(sp);
    if (rc) return rc;
    if (nframes < 0 || (!frames && nframes > 0)) return fail(QCS_ERR_ARG, "invalid frame buffer");
    const size_t npx = (size_t)sp->width * sp->height;
    std::vector<double> img(npx);
    for (int f = 0; f < nframes; f++) {
        const int frame = first_frame + f;
        const int nem = qcs_sim::frame_slots(*sp, seed, frame);
        std::fill(img.begin(), img.end(), 0.0);
        for (int k = 0; k < nem; k++) {
            Emitter e;
            if (!qcs_sim::frame_slot_emitter(*sp, seed, frame, k, e)) continue;
            const double reach = qcs_sim::emitter_reach(e);
            int c0 = (int)ceil(e.x - reach - 0.5), c1 = (int)floor(e.x + reach - 0.5);
            int r0 = (int)ceil(e.y - reach - 0.5), r1 = (int)floor(e.y + reach - 0.5);
            c0 = c0 < 0 ? 0 : c0;
            r0 = r0 < 0 ? 0 : r0;
            c1 = c1 > sp->width - 1 ? sp->width - 1 : c1;
            r1 = r1 > sp->height - 1 ? sp->height - 1 : r1;
            for (int r = r0; r <= r1; r++)
                for (int c = c0; c <= c1; c++) img[(size_t)r * sp->width + c] += qcs_sim::psf_value(e, c + 0.5, r + 0.5);
        }
        const double bg = qcs_sim::frame_background(*sp, seed, frame);
        uint16_t *out = frames + (size_t)f * npx;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)npx; i++) {
            qcs_sim::Gen g(seed, (uint32_t)frame, (uint32_t)i, qcs_sim::TAG_FRAME_PIX);
            const int e = qcs_sim::poisson((double)sp->qe * (bg + img[i]), g);
            out[i] = qcs_sim::to_adu(e, *sp);
        }
    }
    return QCS_OK;
}

int qcs_sim_frames_dev(const qcs_sim_params *sp, uint64_t seed, int first_frame, int nframes, uint16_t *d_frames,
                       void *stream)
{
    int rc = check_sim(sp);
    if (rc) return rc;
    if (nframes <= 0) return QCS_OK;
    hipStream_t s = (hipStream_t)stream;
    const size_t npx = (size_t)sp->width * sp->height;
    double *img = nullptr;
    QCS_HIP(hipMallocAsync((void **)&img, npx * sizeof(double), s));
    for (int f = 0; f < nframes; f++) {
        const int frame = first_frame + f;
        const int nem = qcs_sim::frame_slots(*sp, seed, frame);
        hipError_t e = hipMemsetAsync(img, 0, npx * sizeof(double), s);
        if (e == hipSuccess && nem > 0) {
            hipLaunchKernelGGL(sim_splat_kernel, dim3((nem + 255) / 256), dim3(256), 0, s, *sp, seed, frame, nem, img);
            e = hipGetLastError();
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(sim_noise_kernel, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, s, *sp, seed, frame, img,
                               d_frames + (size_t)f * npx);
            e = hipGetLastError();
        }
        if (e != hipSuccess) {
            (void)hipFreeAsync(img, s);
            return hip_fail(e, "frame simulation", __FILE__, __LINE__);
        }
    }
    QCS_HIP(hipFreeAsync(img, s));
    return QCS_OK;
}

}  // extern "C"
