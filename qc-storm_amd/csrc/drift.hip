// drift.hip -- drift correction by cross-correlation of super-resolution images on gfx950.
//
// Reference (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll/DriftCorrection_SRImageXCorr):
//   CorrectSampleShift                 DLL/DriftCorrection_SRImageXCorr_top.cu:18-74
//   GetSliceShift                      DLL/DriftCorrection_SRImageXCorr_top.cu:172-298
//   ShiftInterpolation, CalcGroupNum, GetCorrGroupFluoPos, GetAFrameEndFluoPos, GetTotalFrame
//                                      DLL/DriftCorrection_SRImageXCorr_rc.cu:20-73, 170-287, 360-395
//   PointFillRegionWithPrec1, gpuApplyShiftCorr, gpuCrossCorr_Mul/_SumOfMul
//                                      DLL/DriftCorrection_SRImageXCorr.cu:21-212
//
// The frames are split into groups of corr_frames overlapping by half; each group's rows are
// rendered at 8x zoom (unit-weight Gaussians of the CRLB width, 7x7 SR pixels); every group is
// cross-correlated with group 0 over a 47 x 47 window of SR shifts centred on the previous group's
// integer shift; the centre of mass of the correlation peak gives the group's shift, and the shifts
// are interpolated linearly per frame and subtracted from X and Y.
//
// The reference evaluates each of the 2209 shifts as its own pass over the whole SR image
// (multiply kernel, column-sum kernel, host sum).  Here the 47 x 47 window is one GEMM per image
// pair:  CC[sx][sy] = sum over img2 rows r and img1 columns x of
//          img2[r][x + bx + sx] * img1[r - by - sy][x]
// (out-of-image pixels read as 0, which reproduces the reference's in-bounds products); the first
// factor is a Hankel matrix of img2 rows (depends on sx), the second a Toeplitz matrix of img1 rows
// (depends on sy).  Work-groups stage a 16-row x 64-column tile of both images in LDS and run FP32
// MFMA 16x16x4 on the 48 x 48 (padded) window; each of the 4 waves keeps nine 16 x 16 accumulator
// blocks, the waves are summed in LDS in a fixed order and each work-group adds its window to an
// FP64 accumulator.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <algorithm>
#include <vector>
#include "qcs_devmath.h"
#include "qcs_internal.h"

namespace qcs {

namespace {

constexpr int CORR = 47, HALF = 23, FITR = 12, DZOOM = 8, RW = 7, RW_HALF = 3;
constexpr int WIN = 48;                 // padded window (3 x 16)
constexpr int TR = 16, TX = 64;         // tile: img2 rows x img1 columns
constexpr int P2 = TX + WIN;            // img2 tile pitch (columns x .. x + 47 of the window)
constexpr int R1 = TR + WIN - 1;        // img1 tile rows
constexpr int P1 = TX + 1;              // img1 tile pitch (odd: the B-operand reads walk rows)

typedef float f32x4_t __attribute__((ext_vector_type(4)));

__global__ void drift_fill_kernel(float *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0f;
}

// the frame column of the rows, contiguous (one copy to the host instead of a strided one)
__global__ void drift_frames_kernel(const float *__restrict__ loc, long long n, float *__restrict__ fr)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        fr[i] = loc[(size_t)i * QCS_REC_FLOATS + QCS_POS_FRME];
}

__global__ void drift_fill64_kernel(double *p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

// PointFillRegionWithPrec1 (DriftCorrection_SRImageXCorr.cu:48-125) at zoom 8
__global__ void __launch_bounds__(256) drift_render_kernel(const float *__restrict__ loc, long long n, int srw, int srh,
                                                           float px_sr, float snr_th, float *__restrict__ img)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *r = loc + (size_t)i * QCS_REC_FLOATS;
    const float pp = r[QCS_POS_PPHO];
    const float rx = r[QCS_POS_XPOS] * (float)DZOOM, ry = r[QCS_POS_YPOS] * (float)DZOOM;
    const float snr = r[QCS_POS_PSNR];
    float prec = r[QCS_POS_CRBX];
    if (!((snr >= snr_th) && (pp > 1.0f) && (rx > (float)RW) && (rx < (float)(srw - 1 - RW)) && (ry > (float)RW) &&
          (ry < (float)(srh - 1 - RW))))
        return;
    if (!((prec > 0.01f) && (prec < 35.0f))) prec = 35.0f;
    const float sig = prec / px_sr;
    const float rs = -0.5f / (sig * sig);
    const float fx = (rx - (float)(int)rx) + (float)RW_HALF, fy = (ry - (float)(int)ry) + (float)RW_HALF;
    for (int yy = 0; yy < RW; yy++) {
        float *row = img + (size_t)srw * (size_t)((int)ry - RW_HALF + yy) + (size_t)((int)rx - RW_HALF);
        const float dy = (float)yy - fy;
#pragma unroll
        for (int xx = 0; xx < RW; xx++) {
            const float dx = (float)xx - fx;
            atomicAdd(row + xx, 1.0f * qexpf(__builtin_fmaf(dx, dx, dy * dy) * rs));
        }
    }
}

// The 48 x 48 correlation window of img1 and img2 at bias (bx, by); acc[sy][sx] (FP64, 48 x 48)
__global__ void __launch_bounds__(256) drift_cc_kernel(const float *__restrict__ img1, const float *__restrict__ img2, int w,
                                                       int h, int bx, int by, int nchunk_x, int nchunks,
                                                       double *__restrict__ acc)
{
    __shared__ float s2[TR * P2];
    __shared__ float s1[R1 * P1];
    __shared__ float red[WIN * WIN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    f32x4_t c[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) c[a][b] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};

    for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const int r0 = (ch / nchunk_x) * TR, x0 = (ch % nchunk_x) * TX;
        // img2 rows r0 .. r0 + 15, columns x0 + bx - 23 .. + P2 - 1
        for (int k = tid; k < TR * P2; k += 256) {
            const int rr = r0 + k / P2, cc = x0 + bx - HALF + k % P2;
            s2[k] = (rr < h && cc >= 0 && cc < w) ? img2[(size_t)rr * w + cc] : 0.0f;
        }
        // img1 rows r0 - by - 24 .. + R1 - 1, columns x0 .. x0 + 63
        for (int k = tid; k < R1 * TX; k += 256) {
            const int rr = r0 - by - (HALF + 1) + k / TX, cc = x0 + k % TX;
            s1[(k / TX) * P1 + k % TX] = (rr >= 0 && rr < h && cc < w) ? img1[(size_t)rr * w + cc] : 0.0f;
        }
        __syncthreads();
        for (int rl = wave * (TR / 4); rl < (wave + 1) * (TR / 4); rl++) {
#pragma unroll 4
            for (int q = 0; q < TX / 4; q++) {
                const int xl = 4 * q + lk;
                float av[3], bv[3];
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    av[b] = s2[rl * P2 + xl + 16 * b + li];                       // img2[r][x + bx + sx]
                    bv[b] = s1[(rl - 16 * b - li + WIN - 1) * P1 + xl];           // img1[r - by - sy][x]
                }
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++) c[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], c[a][b], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // sum the waves in a fixed order: D[4 (lane / 16) + v][lane % 16] of block (a, b) is
    // CC[sx = 16 a + 4 (lane / 16) + v][sy = 16 b + lane % 16]
    for (int wv = 0; wv < 4; wv++) {
        if (wave == wv) {
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int sx = 16 * a + 4 * lk + v, sy = 16 * b + li;
                        float &d = red[sy * WIN + sx];
                        d = (wv == 0) ? c[a][b][v] : d + c[a][b][v];
                    }
        }
        __syncthreads();
    }
    for (int k = tid; k < WIN * WIN; k += 256) {
        const int sy = k / WIN, sx = k % WIN;
        if (sx < CORR && sy < CORR) atomicAdd(&acc[k], (double)red[k]);
    }
}

// gpuApplyShiftCorr (DriftCorrection_SRImageXCorr.cu:127-157); rows of frames outside
// [1, total] are left as they are
__global__ void __launch_bounds__(256) drift_apply_kernel(float *__restrict__ loc, long long n, const float *__restrict__ fx,
                                                          const float *__restrict__ fy, int total)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float *r = loc + (size_t)i * QCS_REC_FLOATS;
    const int f = (int)r[QCS_POS_FRME];
    if (f < 1 || f > total) return;
    r[QCS_POS_XPOS] = r[QCS_POS_XPOS] - fx[f - 1];
    r[QCS_POS_YPOS] = r[QCS_POS_YPOS] - fy[f - 1];
}

// ---- host side: the reference's group bookkeeping and peak refinement ----

int group_count(int total, int corr, int *gs, int *ge)
{
    const int half = corr / 2;
    int g = (total / corr) * 2 - 1;
    if (g <= 0 || g > QCS_DRIFT_MAX_GROUPS - 1) return -1;
    const int resid = total % corr;
    for (int k = 0; k < g; k++) {
        gs[k] = k * half + 1;
        ge[k] = k * half + corr;
    }
    if (resid < half) {
        ge[g - 1] = total;
    } else {
        g++;
        gs[g - 1] = total - corr + 1;
        ge[g - 1] = total;
    }
    return g;
}

// GetAFrameEndFluoPos: row before the first row at/after `from` with a frame > 0 and > find
long long end_row(const std::vector<float> &fr, long long from, int find)
{
    const long long n = (long long)fr.size();
    long long e = 0;
    for (long long i = from; i < n; i++) {
        const int f = (int)fr[(size_t)i];
        if (f > 0 && f > find) {
            e = i - 1;
            break;
        }
    }
    return e == 0 ? n - 1 : e;
}

// GetSliceShift after the correlations: first maximum, normalisation, centre of mass
void peak_shift(const double *cc, int bx, int by, float *sx_out, float *sy_out)
{
    double c[CORR * CORR];
    double mx = 0.0;
    int px = 0, py = 0;
    for (int y = 0; y < CORR; y++)
        for (int x = 0; x < CORR; x++) {
            c[y * CORR + x] = cc[y * WIN + x];
            if (mx < c[y * CORR + x]) {
                mx = c[y * CORR + x];
                px = x;
                py = y;
            }
        }
    for (int k = 0; k < CORR * CORR; k++) c[k] = c[k] / mx * 1000;
    const int xs = std::max(px - FITR, 0), ys = std::max(py - FITR, 0);
    const int xe = std::min(px + FITR, CORR - 1), ye = std::min(py + FITR, CORR - 1);
    const int xh = (xe - xs + 1) / 2, yh = (ye - ys + 1) / 2;
    float sumx[2 * FITR + 1] = {0}, sumy[2 * FITR + 1] = {0};
    for (int y = ys; y <= ye; y++)
        for (int x = xs; x <= xe; x++) {
            float &ax = sumx[FITR - xh + (x - xs)];
            float &ay = sumy[FITR - yh + (y - ys)];
            ax = (float)((double)ax + c[y * CORR + x]);
            ay = (float)((double)ay + c[y * CORR + x]);
        }
    auto com = [](const float *s) {
        float ws = 0.0f, cs = 0.0f;
        for (int k = 0; k < 2 * FITR + 1; k++) {
            if (s[k] == 0) continue;
            ws += s[k] * (float)(k + 1);
            cs += s[k];
        }
        return ws / cs - (float)(FITR + 1);
    };
    *sx_out = (float)(bx + px - HALF) + com(sumx);
    *sy_out = (float)(by + py - HALF) + com(sumy);
}

// ShiftInterpolation: per-frame shifts (index frame - 1)
void interpolate(int total, int ng, const int *gs, const int *ge, const float *gx, const float *gy, float *fx, float *fy)
{
    std::fill(fx, fx + total, 0.0f);
    std::fill(fy, fy + total, 0.0f);
    float lc = 0, lx = 0, ly = 0, cc = 0, cx = 0, cy = 0, kx = 0, ky = 0;
    for (int k = 0; k <= ng; k++) {
        if (k == ng) {
            cc = (float)ge[ng - 1];
        } else {
            cc = (float)((gs[k] + ge[k]) / 2);
            cx = gx[k];
            cy = gy[k];
            kx = (cx - lx) / (cc - lc);
            ky = (cy - ly) / (cc - lc);
        }
        for (int f = (int)lc; (float)f < cc; f++) {
            fx[f] = lx + ((float)f - lc) * kx;
            fy[f] = ly + ((float)f - lc) * ky;
        }
        lc = cc;
        lx = cx;
        ly = cy;
    }
}

}  // namespace

}  // namespace qcs

using namespace qcs;

extern "C" {

int qcs_drift_correct_dev(qcs_ctx *ctx, float *d_rows, long long n, int corr_frames, float *shift_x, float *shift_y,
                          int shift_cap, int *total_frames, float *group_shift_xy, void *stream)
{
    if (!ctx || n < 0 || (n > 0 && !d_rows) || corr_frames < 2 || ((shift_x || shift_y) && shift_cap < 0))
        return fail(QCS_ERR_ARG, "invalid argument");
    if (total_frames) *total_frames = 0;
    if (n == 0) return fail(QCS_ERR_ARG, "no rows");
    const qcs_params &p = ctx->p;
    if (p.ImageWidth < 2 * RW || p.ImageHigh < 2 * RW) return fail(QCS_ERR_ARG, "ImageWidth/ImageHigh not set");
    hipStream_t s = (hipStream_t)stream;
    // frame column (GetTotalFrame reads the last 400 rows; the group row ranges need all of it)
    std::vector<float> fr((size_t)n);
    int rc;
    if ((rc = ctx->dr_frames.ensure((size_t)n * sizeof(float)))) return rc;
    hipLaunchKernelGGL(drift_frames_kernel, dim3(1024), dim3(256), 0, s, d_rows, n, ctx->dr_frames.as<float>());
    QCS_HIP(hipGetLastError());
    QCS_HIP(hipMemcpyAsync(fr.data(), ctx->dr_frames.ptr, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipStreamSynchronize(s));
    int total = 0;
    for (long long i = n - 1; i >= n - std::min<long long>(n, 400); i--)
        if ((int)fr[(size_t)i] > 0) {
            total = (int)fr[(size_t)i];
            break;
        }
    if (total_frames) *total_frames = total;
    if ((shift_x || shift_y) && shift_cap < total) return fail(QCS_ERR_CAPACITY, "shift_cap (%d) < frames (%d)", shift_cap, total);
    int gs[QCS_DRIFT_MAX_GROUPS], ge[QCS_DRIFT_MAX_GROUPS];
    const int ng = group_count(total, corr_frames, gs, ge);
    if (ng <= 0)
        return fail(QCS_ERR_ARG, "drift correction needs 1 <= frames / corr_frames and at most %d groups (frames %d, corr %d)",
                    QCS_DRIFT_MAX_GROUPS, total, corr_frames);
    long long first[QCS_DRIFT_MAX_GROUPS], last[QCS_DRIFT_MAX_GROUPS];
    {
        long long b = 0, e = 0;
        first[0] = 0;
        for (int k = 0; k < ng; k++) {
            e = end_row(fr, e, ge[k]);
            last[k] = e;
            if (k >= 1) {
                b = end_row(fr, b, gs[k] - 1);
                first[k] = b + 1;
            }
        }
    }
    const int srw = p.ImageWidth * DZOOM, srh = p.ImageHigh * DZOOM;
    const size_t npx = (size_t)srw * srh;
    if ((rc = ctx->dr_img.ensure(2 * npx * sizeof(float))) || (rc = ctx->dr_cc.ensure(WIN * WIN * sizeof(double))) ||
        (rc = ctx->dr_shift.ensure((size_t)2 * total * sizeof(float))))
        return rc;
    float *img1 = ctx->dr_img.as<float>(), *img2 = img1 + npx;
    double *acc = ctx->dr_cc.as<double>();
    const int nchunk_x = (srw + TX - 1) / TX, nchunks = nchunk_x * ((srh + TR - 1) / TR);
    const int cc_grid = std::min(nchunks, 2048);
    std::vector<double> cc(WIN * WIN);
    std::vector<float> gx((size_t)ng), gy((size_t)ng);
    float sxv = 0.0f, syv = 0.0f;
    int bx = 0, by = 0;
    for (int k = 0; k < ng; k++) {
        float *im = k == 0 ? img1 : img2;
        hipLaunchKernelGGL(drift_fill_kernel, dim3(2048), dim3(256), 0, s, im, npx);
        const long long cnt = last[k] - first[k] + 1;
        if (cnt > 0)
            hipLaunchKernelGGL(drift_render_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
                               d_rows + (size_t)first[k] * QCS_REC_FLOATS, cnt, srw, srh, p.PixelSize / (float)DZOOM, p.SNR_th, im);
        QCS_HIP(hipGetLastError());
        if (k > 0) {
            hipLaunchKernelGGL(drift_fill64_kernel, dim3(1), dim3(256), 0, s, acc, WIN * WIN);
            hipEvent_t ev;
            if ((rc = timer_begin(ctx->timer, s, &ev))) return rc;
            hipLaunchKernelGGL(drift_cc_kernel, dim3(cc_grid), dim3(256), 0, s, img1, img2, srw, srh, bx, by, nchunk_x, nchunks, acc);
            QCS_HIP(hipGetLastError());
            if ((rc = timer_end(ctx->timer, TM_DRIFTCC, s, ev, (long long)npx))) return rc;
            QCS_HIP(hipMemcpyAsync(cc.data(), acc, WIN * WIN * sizeof(double), hipMemcpyDeviceToHost, s));
            QCS_HIP(hipStreamSynchronize(s));
            // cc holds acc[sy][sx]; the reference's CorrResult[y][x] is shift (x - 23, y - 23)
            peak_shift(cc.data(), bx, by, &sxv, &syv);
            bx = (int)sxv;
            by = (int)syv;
        }
        gx[(size_t)k] = sxv / (float)DZOOM;
        gy[(size_t)k] = syv / (float)DZOOM;
    }
    std::vector<float> fxy((size_t)2 * total);
    interpolate(total, ng, gs, ge, gx.data(), gy.data(), fxy.data(), fxy.data() + total);
    float *d_fx = ctx->dr_shift.as<float>();
    QCS_HIP(hipMemcpyAsync(d_fx, fxy.data(), (size_t)2 * total * sizeof(float), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(drift_apply_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_rows, n, d_fx, d_fx + total,
                       total);
    QCS_HIP(hipGetLastError());
    QCS_HIP(hipStreamSynchronize(s));
    if (shift_x) memcpy(shift_x, fxy.data(), (size_t)total * sizeof(float));
    if (shift_y) memcpy(shift_y, fxy.data() + total, (size_t)total * sizeof(float));
    if (group_shift_xy)
        for (int k = 0; k < ng; k++) {
            group_shift_xy[2 * k] = gx[(size_t)k];
            group_shift_xy[2 * k + 1] = gy[(size_t)k];
        }
    return QCS_OK;
}

}  // extern "C"
