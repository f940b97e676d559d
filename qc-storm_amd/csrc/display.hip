// display.hip -- the colour-coded 3D super-resolution image (lm_GetSMLMImage3D) on gfx950.
//
// Reference (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll/ImgRender):
//   RenderSaveImg_3D, ColorGenerate, HSVtoRGB          DLL/ImageRender_3d.cu:238-346, 424-520
//   ImgRender_GetSaveImg_3D                            DLL/ImageRender_3d.cu:375-403
//   gpuGetSRImageHistgram, GetSRImageHistgram          DLL/ImageRender_3d.cu:529-590
//   encodings and colour modes                         DLL/ImageRender_base.h:43-57
//
// The brightness scale is the intensity below which 97% of the image's histogrammed pixels lie
// (2048 bins over [0, 12000)); each pixel's hue comes from its intensity-weighted depth
// (z plane / intensity plane, blue to red or red to blue over [MinZDepth, MaxZDepth]) and its
// value from intensity * 120 / scale, capped at 100.  The histogram counts in LDS per work-group
// (integer, exact); the colour pass is one thread per pixel with coalesced reads and writes.
#include <limits.h>
#include <string.h>
#include <vector>
#include "qcs_internal.h"

namespace qcs {

namespace {

constexpr int HN = 2048;
constexpr float HMAX = 12000.0f;

__device__ __forceinline__ int f2i_sat(float x)
{
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT_MAX;
    if (x <= -2147483648.0f) return INT_MIN;
    return (int)x;
}

__global__ void __launch_bounds__(256) disp_hist_kernel(const float *__restrict__ img, size_t npx, int *__restrict__ hist)
{
    __shared__ int h[HN];
    for (int k = threadIdx.x; k < HN; k += blockDim.x) h[k] = 0;
    __syncthreads();
    const float gap = HMAX / HN;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx; i += (size_t)gridDim.x * blockDim.x) {
        const int pos = f2i_sat(img[i] / gap - 1);
        if (pos >= 0 && pos < HN) atomicAdd(&h[pos], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < HN; k += blockDim.x)
        if (h[k]) atomicAdd(&hist[k], h[k]);
}

__global__ void disp_clear_kernel(int *hist)
{
    for (int k = threadIdx.x; k < HN; k += blockDim.x) hist[k] = 0;
}

__device__ __forceinline__ void hsv_rgb(int h, int s, int v, int &r, int &g, int &b)
{
    h = h < 0 ? 0 : h;
    s = s < 0 ? 0 : s;
    v = v < 0 ? 0 : v;
    const float mx = (float)v * 2.55f;
    const float mn = mx * (float)(100 - s) / 100.0f;
    const int i = h / 60, d = h % 60;
    const float adj = (mx - mn) * (float)d / 60.0f;
    switch (i) {
    case 0: r = (int)(mx + 0.5f); g = (int)(mn + adj + 0.5f); b = (int)(mn + 0.5f); break;
    case 1: r = (int)(mx - adj + 0.5f); g = (int)(mx + 0.5f); b = (int)(mn + 0.5f); break;
    case 2: r = (int)(mn + 0.5f); g = (int)(mx + 0.5f); b = (int)(mn + adj + 0.5f); break;
    case 3: r = (int)(mn + 0.5f); g = (int)(mx - adj + 0.5f); b = (int)(mx + 0.5f); break;
    case 4: r = (int)(mn + adj + 0.5f); g = (int)(mn + 0.5f); b = (int)(mx + 0.5f); break;
    default: r = (int)(mx + 0.5f); g = (int)(mn + 0.5f); b = (int)(mx - adj + 0.5f); break;
    }
    r = min(max(r, 0), 255);
    g = min(max(g, 0), 255);
    b = min(max(b, 0), 255);
}

__global__ void __launch_bounds__(256) disp_rgb_kernel(const float *__restrict__ img, const float *__restrict__ zimg, size_t npx,
                                                       float scale, float zmin, float zmax, int color_mode, int mode,
                                                       uint8_t *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    const int nb = mode < 2 ? 3 : 4;
    const float it = img[i];
    int r = 0, g = 0, b = 0;
    if (it != 0.0f) {
        float z = zimg[i] / it;
        int v = f2i_sat(it * 120.0f / scale);
        v = v >= 100 ? 100 : v;
        z = z < zmin ? zmin : z;
        z = z > zmax ? zmax : z;
        int h = f2i_sat((z - zmin) * 240.0f / (zmax - zmin));
        if (color_mode == 0) h = 240 - h;
        h = h < 0 ? 0 : (h > 240 ? 240 : h);
        hsv_rgb(h, 100, v, r, g, b);
    }
    uint8_t *o = out + i * (size_t)nb;
    if (mode == QCS_RGB_BGR || mode == QCS_RGB_BGRA) {
        o[0] = (uint8_t)b; o[1] = (uint8_t)g; o[2] = (uint8_t)r;
    } else {
        o[0] = (uint8_t)r; o[1] = (uint8_t)g; o[2] = (uint8_t)b;
    }
    if (nb == 4) o[3] = 0xff;
}

// GetSRImageHistgram's 97% level from the counts, summed in float as the reference does
float level97(const int *h)
{
    const float gap = HMAX / HN;
    float tot = 0.0f, cur = 0.0f, gi = (float)(HN - 1);
    for (int k = 0; k < HN; k++) tot += (float)h[k];
    for (int k = 0; k < HN; k++) {
        cur += (float)h[k];
        if (cur / tot > 0.97f) {
            gi = (float)k;
            break;
        }
    }
    return gi * gap;
}

}  // namespace

int render_rgb_launch(qcs_ctx *ctx, const float *d_img, const float *d_zimg, int mode, float bright, uint8_t *d_out,
                      float *max_intensity, hipStream_t s)
{
    const qcs_params &p = ctx->p;
    const size_t npx = (size_t)p.SRImageWidth * p.SRImageHigh;
    int rc;
    if ((rc = ctx->dp_hist.ensure(HN * sizeof(int)))) return rc;
    int *hist = ctx->dp_hist.as<int>();
    hipLaunchKernelGGL(disp_clear_kernel, dim3(1), dim3(256), 0, s, hist);
    const unsigned blocks = (unsigned)std::min<size_t>((npx + 255) / 256, 1024);
    hipLaunchKernelGGL(disp_hist_kernel, dim3(blocks), dim3(256), 0, s, d_img, npx, hist);
    QCS_HIP(hipGetLastError());
    std::vector<int> h(HN);
    QCS_HIP(hipMemcpyAsync(h.data(), hist, HN * sizeof(int), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipStreamSynchronize(s));
    const float scale = level97(h.data()) * bright;
    if (max_intensity) *max_intensity = scale;
    hipLaunchKernelGGL(disp_rgb_kernel, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, s, d_img, d_zimg, npx, scale,
                       p.MinZDepth, p.MaxZDepth, p.ColorMode_3D, mode, d_out);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

}  // namespace qcs

using namespace qcs;

extern "C" int qcs_render_rgb_dev(qcs_ctx *ctx, const float *d_img, const float *d_zimg, int mode, float bright_ratio,
                                  uint8_t *d_out, float *max_intensity, void *stream)
{
    if (!ctx || !d_img || !d_zimg || !d_out || mode < QCS_RGB_BGR || mode > QCS_RGB_RGBA || !(bright_ratio > 0.0f))
        return fail(QCS_ERR_ARG, "invalid argument");
    const qcs_params &p = ctx->p;
    if (p.SRImageWidth <= 0 || p.SRImageHigh <= 0) return fail(QCS_ERR_ARG, "SRImageWidth/SRImageHigh not set");
    return render_rgb_launch(ctx, d_img, d_zimg, mode, bright_ratio, d_out, max_intensity, (hipStream_t)stream);
}
