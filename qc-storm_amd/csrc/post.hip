// post.hip -- frame-sorted compaction of localization rows and super-resolution splat rendering.
//
// Reference (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll):
//   RemoveZero_SortFrame   DLL/bfgsMLE_Loc_top/ZeroLocalizationsRemove.cu:42-61,161-186
//   RemoveZeroWithoutSort  DLL/bfgsMLE_Loc_top/ZeroLocalizationsRemove.cu:64-102,142-159
//   CreateRenderPSF        DLL/ImgRender/ImageRender_2d.cu:25-47
//   FluoRenderWithPrec_2D  DLL/ImgRender/ImageRender_2d.cu:55-155
//   FluoRenderWithPrec_3D  DLL/ImgRender/ImageRender_3d.cu:60-147
//
// Sorted compaction: the reference launches one kernel per frame and appends with an atomic
// counter (within-frame order arbitrary, O(frames x rows) work).  Here the kept rows are compacted
// in row order (scan), then stably radix-sorted by frame (hipCUB), so rows of one frame keep their
// original order.  Rendering: one lane per localization splats a 7 x 7 Gaussian with float
// atomics into the SR image, as the reference does; the PSF uses the deterministic exponential.
#include <hipcub/hipcub.hpp>
#include <string.h>
#include "qcs_internal.h"
#include "qcs_devmath.h"

namespace qcs {

// float -> int truncation with NaN -> 0 and saturation (the CUDA conversion the reference relies on)
__device__ __forceinline__ int f2i_sat(float v)
{
    if (v != v) return 0;
    if (v >= 2147483647.0f) return 2147483647;
    if (v <= -2147483648.0f) return (-2147483647 - 1);
    return (int)v;
}

__global__ void __launch_bounds__(256) keep_sorted_kernel(const float *__restrict__ loc, int n, int first, int end,
                                                          int *__restrict__ flag, unsigned *__restrict__ key)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *r = loc + (size_t)i * QCS_REC_FLOATS;
    const int pp = f2i_sat(r[QCS_POS_PPHO]), fr = f2i_sat(r[QCS_POS_FRME]);
    const bool k = ((float)pp > 1.0f) && fr >= first && fr <= end;
    flag[i] = k ? 1 : 0;
    key[i] = k ? (unsigned)(fr - first) : 0u;
}

__global__ void __launch_bounds__(256) keep_any_kernel(const float *__restrict__ loc, int n, int *__restrict__ flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (loc[(size_t)i * QCS_REC_FLOATS + QCS_POS_PPHO] > 1.0f) ? 1 : 0;
}

// gather kept rows: keys_out[pos] = key, idx_out[pos] = row
__global__ void __launch_bounds__(256) gather_keys_kernel(const int *__restrict__ flag, const int *__restrict__ offs,
                                                          const unsigned *__restrict__ key, int n,
                                                          unsigned *__restrict__ kout, int *__restrict__ iout)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    kout[offs[i]] = key[i];
    iout[offs[i]] = i;
}

__global__ void __launch_bounds__(256) copy_rows_kernel(const float *__restrict__ loc, const int *__restrict__ idx,
                                                        const int *__restrict__ m_dev, float *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= *m_dev) return;
    const float4 *src = reinterpret_cast<const float4 *>(loc + (size_t)idx[j] * QCS_REC_FLOATS);
    float4 *dst = reinterpret_cast<float4 *>(out + (size_t)j * QCS_REC_FLOATS);
    dst[0] = src[0];
    dst[1] = src[1];
    dst[2] = src[2];
}

// ---------------------------------------------------------------------------------------------
struct RenderConst {
    int srw, srh, mode, is3d;
    float zoom, ps, snr_th, fixed_prec, zth;
};

__global__ void __launch_bounds__(256) render_kernel(const float *__restrict__ loc, int n, RenderConst c,
                                                     float *__restrict__ img, float *__restrict__ zimg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *r = loc + (size_t)i * QCS_REC_FLOATS;
    float pp = r[QCS_POS_PPHO];
    const float rx = r[QCS_POS_XPOS] * c.zoom, ry = r[QCS_POS_YPOS] * c.zoom;
    float snr = r[QCS_POS_PSNR];
    float lp = c.is3d ? (r[QCS_POS_CRBX] + r[QCS_POS_CRBY]) / 2.0f : r[QCS_POS_CRBX];
    const float z = r[QCS_POS_ZPOS];
    if (lp != lp) { snr = 0.0f; pp = 0.0f; }
    if (lp <= 0.0f) pp = 0.0f;
    bool inside = snr >= c.snr_th && pp > 1.0f && rx > 7.0f && rx < (float)(c.srw - 1 - 7) && ry > 7.0f &&
                  ry < (float)(c.srh - 1 - 7);
    if (c.is3d) inside = inside && fabsf(z) < c.zth;
    if (!inside) return;
    if (c.mode == QCS_RENDER_1PHOTON_CRLB || c.mode == QCS_RENDER_1PHOTON_FIXED) pp = 1.0f;
    if (c.mode == QCS_RENDER_1PHOTON_FIXED) lp = c.fixed_prec;
    if (!c.is3d && lp > 30.0f) { lp = 30.0f; pp = 0.0f; }
    lp = fmaxf(lp, (c.ps / c.zoom) / 2.0f);
    float rs = lp / (c.ps / c.zoom);
    rs = -0.5f / (rs * rs);
    const int ix = (int)rx, iy = (int)ry;
    const float fx = (rx - (float)ix) + 3.0f, fy = (ry - (float)iy) + 3.0f;
    if (!c.is3d && c.mode == QCS_RENDER_PEAKPHOTON_1PIXEL) {
        atomicAdd(&img[(size_t)c.srw * iy + ix], pp);
        return;
    }
#pragma unroll
    for (int a = 0; a < 7; a++) {
        const size_t off = (size_t)c.srw * (iy - 3 + a) + ix - 3;
        const float dy = (float)a - fy;
#pragma unroll
        for (int b = 0; b < 7; b++) {
            const float dx = (float)b - fx;
            const float psf = qexpf(__builtin_fmaf(dx, dx, dy * dy) * rs);
            atomicAdd(&img[off + b], pp * psf);
            if (c.is3d && zimg) atomicAdd(&zimg[off + b], (z * pp) * psf);
        }
    }
}

__global__ void __launch_bounds__(256) max_disp_kernel(const float *__restrict__ img, size_t n, int *__restrict__ out)
{
    int m = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)img[i];
        m = v > m ? v : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_down(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

int render_launch(const qcs_params &p, const float *d_loc, int n, float *d_img, float *d_zimg, hipStream_t s)
{
    if (n <= 0) return QCS_OK;
    RenderConst c;
    c.srw = p.SRImageWidth;
    c.srh = p.SRImageHigh;
    c.mode = p.RenderingMode;
    c.is3d = (p.LocType == QCS_LOCTYPE_AS3D || p.LocType == QCS_LOCTYPE_DH3D);
    c.zoom = p.PixelZoom;
    c.ps = p.PixelSize;
    c.snr_th = p.SNR_th;
    c.fixed_prec = p.FixedLocPrec;
    c.zth = 5.0f * fmaxf(fabsf(p.MinZDepth), fabsf(p.MaxZDepth));
    hipLaunchKernelGGL(render_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_loc, n, c, d_img, d_zimg);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

int max_disp_launch(const float *d_img, size_t npx, int *d_out, hipStream_t s)
{
    size_t blocks = (npx + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(max_disp_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d_img, npx, d_out);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

}  // namespace qcs

using namespace qcs;

extern "C" int qcs_compact_sorted_dev(qcs_ctx *c, const float *d_loc, int n, int first, int end, float *d_out, int *n_out,
                                      void *stream)
{
    if (n_out) *n_out = 0;
    if (!c || n < 0 || (n > 0 && (!d_loc || !d_out))) return fail(QCS_ERR_ARG, "invalid argument");
    if (n == 0 || end < first) return QCS_OK;
    hipStream_t s = (hipStream_t)stream;
    const int ntiles = scan_tile_count((size_t)n);
    // flag, offsets, tile sums, keys in/out, indices
    size_t temp = 0;
    QCS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const unsigned *)nullptr, (unsigned *)nullptr,
                                               (const int *)nullptr, (int *)nullptr, n, 0, 32, s));
    const size_t words = (size_t)n * 5 + ntiles + 1;
    int rc;
    if ((rc = c->fit_flag.ensure(sizeof(int) * words)) || (rc = c->fit_rows.ensure(temp + 256)) ||
        (rc = c->fit_idx1.ensure(sizeof(int) * (size_t)n)) || (rc = c->fit_cnt.ensure(64)) || (rc = c->pin_cnt.ensure(64)))
        return rc;
    int *flag = c->fit_flag.as<int>(), *offs = flag + n, *tsum = offs + n;
    unsigned *key = reinterpret_cast<unsigned *>(tsum + ntiles + 1), *kin = key + n, *kout = kin + n;
    int *iin = c->fit_idx1.as<int>();
    int *iout = flag;   // flag is dead after the gather
    const dim3 g((n + 255) / 256);
    hipEvent_t ev;
    if ((rc = timer_begin(c->timer, s, &ev))) return rc;
    hipLaunchKernelGGL(keep_sorted_kernel, g, dim3(256), 0, s, d_loc, n, first, end, flag, key);
    if ((rc = scan_exclusive(flag, n, offs, tsum, s))) return rc;
    hipLaunchKernelGGL(gather_keys_kernel, g, dim3(256), 0, s, flag, offs, key, n, kin, iin);
    QCS_HIP(hipGetLastError());
    int *h = c->pin_cnt.as<int>();
    QCS_HIP(hipMemcpyAsync(h, tsum + ntiles, sizeof(int), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipStreamSynchronize(s));
    const int m = h[0];
    if (m > 0) {
        int bits = 1;
        while (bits < 32 && ((unsigned)(end - first) >> bits) != 0) bits++;
        size_t tb = 0;
        QCS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, kout, iin, iout, m, 0, bits, s));
        if (tb > c->fit_rows.bytes) return fail(QCS_ERR_NOMEM, "radix sort scratch larger than reserved");
        QCS_HIP(hipcub::DeviceRadixSort::SortPairs(c->fit_rows.ptr, tb, kin, kout, iin, iout, m, 0, bits, s));
        QCS_HIP(hipMemcpyAsync(c->fit_cnt.ptr, &h[0], sizeof(int), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(copy_rows_kernel, dim3((m + 255) / 256), dim3(256), 0, s, d_loc, iout, c->fit_cnt.as<int>(), d_out);
        QCS_HIP(hipGetLastError());
    }
    if ((rc = timer_end(c->timer, TM_COMPACT, s, ev, n))) return rc;
    QCS_HIP(hipStreamSynchronize(s));
    if (n_out) *n_out = m;
    return QCS_OK;
}

extern "C" int qcs_compact_dev(qcs_ctx *c, const float *d_loc, int n, float *d_out, int *n_out, void *stream)
{
    if (n_out) *n_out = 0;
    if (!c || n < 0 || (n > 0 && (!d_loc || !d_out))) return fail(QCS_ERR_ARG, "invalid argument");
    if (n == 0) return QCS_OK;
    hipStream_t s = (hipStream_t)stream;
    const int ntiles = scan_tile_count((size_t)n);
    int rc;
    if ((rc = c->fit_flag.ensure(sizeof(int) * ((size_t)n * 2 + ntiles + 1))) || (rc = c->fit_idx1.ensure(sizeof(int) * (size_t)n)) ||
        (rc = c->pin_cnt.ensure(64)))
        return rc;
    int *flag = c->fit_flag.as<int>(), *offs = flag + n, *tsum = offs + n;
    int *idx = c->fit_idx1.as<int>();
    hipLaunchKernelGGL(keep_any_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_loc, n, flag);
    if ((rc = scan_exclusive(flag, n, offs, tsum, s))) return rc;
    if ((rc = compact_scatter(flag, offs, n, nullptr, nullptr, 0, idx, s))) return rc;
    hipLaunchKernelGGL(copy_rows_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_loc, idx, tsum + ntiles, d_out);
    QCS_HIP(hipGetLastError());
    int *h = c->pin_cnt.as<int>();
    QCS_HIP(hipMemcpyAsync(h, tsum + ntiles, sizeof(int), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipStreamSynchronize(s));
    if (n_out) *n_out = h[0];
    return QCS_OK;
}

extern "C" int qcs_render_dev(qcs_ctx *c, const float *d_loc, int n, float *d_img, float *d_zimg, void *stream)
{
    if (!c || n < 0 || (n > 0 && (!d_loc || !d_img))) return fail(QCS_ERR_ARG, "invalid argument");
    if (c->p.SRImageWidth <= 16 || c->p.SRImageHigh <= 16) return fail(QCS_ERR_ARG, "SR image size not set");
    hipEvent_t ev;
    int rc = timer_begin(c->timer, (hipStream_t)stream, &ev);
    if (rc) return rc;
    rc = render_launch(c->p, d_loc, n, d_img, d_zimg, (hipStream_t)stream);
    if (rc) return rc;
    return timer_end(c->timer, TM_RENDER, (hipStream_t)stream, ev, n);
}
