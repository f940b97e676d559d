// wle.hip -- weighted-likelihood (WLE) pre-estimation on gfx950.
//
// Reference (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll):
//   gpu_CalculatePSFWidth       DLL/WLEParaEstimation/WLEParaEstimation_core.h:44-156
//   SplineInterpolation         DLL/WLEParaEstimation/SplineInterpolation.h:40-202
//   FindMaxPos / SigmaWidthCalc DLL/WLEParaEstimation/WLEParaEstimation_core.cu:91-136
//   nearest neighbour           DLL/WLEParaEstimation/WLEParaEstimation_core.cu:140-183
//   molecule type               DLL/WLEParaEstimation/WLEParaEstimation_core.cu:197-264
//
// PSF width: one ROI per lane.  The natural cubic spline through the R row (column) means is
// solved once per axis; its 10(R-1)+1 samples are never stored -- the peak search and the two
// one-sided sums evaluate the spline on the fly with the interval index known at compile time
// (sample j lies in interval j / 10), in the reference's summation order.
// Nearest neighbour: the reference's all-pairs scan (O(N^2)) is kept as the general path; when the
// ROI list is sorted by (frame, y) -- extraction produces it in raster order -- each ROI sweeps
// outward through its frame's segment and stops once the row distance alone exceeds the best
// distance found, which gives the identical minimum (an ROI alone in its frame keeps the scan's
// initial value 1e8, which every other-frame term d^2 + 1e8 and the self term 2e8 cannot beat).
#include "qcs_internal.h"

namespace qcs {

template <int R>
struct Spline {
    float y[R], b[R], c[R], d[R];
};

// natural cubic spline on the knots 0..R-1 (unit spacing)
template <int R>
__device__ __forceinline__ void spline_fit(const float (&yv)[R], Spline<R> &s)
{
    float A_sub[R], A_diag[R], A_sup[R], rhs[R];
#pragma unroll
    for (int i = 0; i < R; i++) { A_sub[i] = 0.0f; A_diag[i] = 0.0f; A_sup[i] = 0.0f; }
    A_diag[0] = 1.0f;
    A_diag[R - 1] = 1.0f;
#pragma unroll
    for (int i = 1; i < R - 1; i++) {
        const float h0 = (float)i - (float)(i - 1), h1 = (float)(i + 1) - (float)i;
        A_sub[i] = h0;
        A_diag[i] = 2.0f * (h0 + h1);
        A_sup[i] = h1;
    }
    rhs[0] = 0.0f;
    rhs[R - 1] = 0.0f;
#pragma unroll
    for (int i = 1; i < R - 1; i++) {
        const float h0 = (float)i - (float)(i - 1), h1 = (float)(i + 1) - (float)i;
        rhs[i] = __builtin_fmaf(3.0f / h1, yv[i + 1] - yv[i], -((3.0f / h0) * (yv[i] - yv[i - 1])));
    }
    float beta[R], yy[R];
#pragma unroll
    for (int i = 0; i < R - 1; i++) {
        if (i == 0) beta[i] = A_sup[i] / A_diag[i];
        else beta[i] = A_sup[i] / __builtin_fmaf(-A_sub[i], beta[i - 1], A_diag[i]);
    }
#pragma unroll
    for (int i = 0; i < R; i++) {
        if (i == 0) yy[i] = rhs[i] / A_diag[i];
        else yy[i] = __builtin_fmaf(-A_sub[i], yy[i - 1], rhs[i]) / __builtin_fmaf(-A_sub[i], beta[i - 1], A_diag[i]);
    }
#pragma unroll
    for (int i = R - 1; i >= 0; i--) {
        if (i == R - 1) s.c[i] = yy[i];
        else s.c[i] = __builtin_fmaf(-beta[i], s.c[i + 1], yy[i]);
    }
#pragma unroll
    for (int i = 0; i < R - 1; i++) {
        const float hi = (float)(i + 1) - (float)i;
        s.b[i] = __builtin_fmaf(1.0f / hi, yv[i + 1] - yv[i], -((hi / 3.0f) * __builtin_fmaf(2.0f, s.c[i], s.c[i + 1])));
        s.d[i] = (s.c[i + 1] - s.c[i]) / (3.0f * hi);
    }
#pragma unroll
    for (int i = 0; i < R; i++) s.y[i] = yv[i];
}

// spline sample j (x = j * 0.1), evaluated as InterplotData does (SplineInterpolation.h:79-101)
template <int R, int J>
__device__ __forceinline__ float spline_at(const Spline<R> &s)
{
    constexpr int PS0 = J / 10;
    constexpr int PS = PS0 >= R - 1 ? R - 2 : PS0;
    const float xn = (float)J * 0.1f;
    const float t = xn - (float)PS;
    float v = __builtin_fmaf(s.b[PS], t, s.y[PS]);
    v = __builtin_fmaf(s.c[PS] * t, t, v);
    v = __builtin_fmaf((s.d[PS] * t) * t, t, v);
    return v;
}

template <int R, int J, int N>
struct SplineScan {
    // peak search over [L, H] (FindMaxPos: strict >, starting at the middle sample)
    __device__ __forceinline__ static void maxpos(const Spline<R> &s, int L, int H, int &mx, float &mv)
    {
        const float v = spline_at<R, J>(s);
        if (J >= L && J <= H && v > mv) { mx = J; mv = v; }
        SplineScan<R, J + 1, N>::maxpos(s, L, H, mx, mv);
    }
    // sums over [mp - cl, mp] and [mp, mp + cl] in ascending order
    __device__ __forceinline__ static void sums(const Spline<R> &s, int lo, int mp, int hi, float &sl, float &sr)
    {
        const float v = spline_at<R, J>(s);
        if (J >= lo && J <= mp) sl = sl + v;
        if (J >= mp && J <= hi) sr = sr + v;
        SplineScan<R, J + 1, N>::sums(s, lo, mp, hi, sl, sr);
    }
};
template <int R, int N>
struct SplineScan<R, N, N> {
    __device__ __forceinline__ static void maxpos(const Spline<R> &, int, int, int &, float &) {}
    __device__ __forceinline__ static void sums(const Spline<R> &, int, int, int, float &, float &) {}
};

template <int R>
__device__ __forceinline__ void two_sided_sigma(const float (&mean)[R], float &sl, float &sr)
{
    constexpr int N = (R - 1) * 10 + 1;
    Spline<R> s;
    spline_fit<R>(mean, s);
    const int ctr = N / 2;
    const int L = ctr - 8 < 0 ? 0 : ctr - 8;
    const int H = ctr + 8 > N - 1 ? N - 1 : ctr + 8;
    // FindMaxPos starts from DataLen/2 and scans [L, H] keeping the first strict maximum
    int mx = ctr;
    float mv = spline_at<R, N / 2>(s);
    SplineScan<R, 0, N>::maxpos(s, L, H, mx, mv);
    const int cl = mx < (N - mx - 1) ? mx : (N - mx - 1);
    float a = 0.0f, b = 0.0f;
    SplineScan<R, 0, N>::sums(s, mx - cl, mx, mx + cl, a, b);
    sl = ((a * 0.1f) / mv) / 1.2533f;
    sr = ((b * 0.1f) / mv) / 1.2533f;
}

template <int R>
__global__ void __launch_bounds__(64) wle_psf_kernel(const uint16_t *__restrict__ rois, int n, const int *n_dev,
                                                     float *__restrict__ wle)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (n_dev) n = *n_dev;
    if (i >= n) return;
    const uint16_t *roi = rois + (size_t)i * (R * (R + 1));
    float mx[R], my[R];
#pragma unroll
    for (int k = 0; k < R; k++) { mx[k] = 0.0f; my[k] = 0.0f; }
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int c = 0; c < R; c++) {
            const float v = (float)roi[r * R + c];
            mx[c] = mx[c] + v;
            my[r] = my[r] + v;
        }
    float mnx = mx[0], mny = my[0];
#pragma unroll
    for (int k = 0; k < R; k++) { mnx = fminf(mnx, mx[k]); mny = fminf(mny, my[k]); }
#pragma unroll
    for (int k = 0; k < R; k++) {
        mx[k] = (mx[k] - mnx) / (float)R;
        my[k] = (my[k] - mny) / (float)R;
    }
    float sl, sr, su, sd;
    two_sided_sigma<R>(mx, sl, sr);
    two_sided_sigma<R>(my, su, sd);
    float *o = wle + (size_t)i * QCS_WLE_FLOATS;
    o[QCS_WLE_SIGMAL] = sl;
    o[QCS_WLE_SIGMAR] = sr;
    o[QCS_WLE_SIGMAU] = su;
    o[QCS_WLE_SIGMAD] = sd;
}

struct RoiKey {
    int x, y, f;
};

__device__ __forceinline__ RoiKey roi_key(const uint16_t *rois, int W, int P, int j)
{
    const uint16_t *h = rois + (size_t)j * W + P;
    RoiKey k;
    k.x = h[0];
    k.y = h[1];
    k.f = (int)h[2] + (int)h[3] * 65536;
    return k;
}

// distance term of the reference (:171-175): integer d^2 plus 1e8 for another frame and for self
__device__ __forceinline__ float nn_term(const RoiKey &a, const RoiKey &b, bool self)
{
    const int bias = (int)((float)(a.f != b.f) * 100000000.0f + (float)self * 100000000.0f);
    return (float)((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) + bias);
}

// all-pairs nearest neighbour (gpu_CalculateNearestNeighborDistance)
__global__ void __launch_bounds__(256) wle_nn_allpairs_kernel(const uint16_t *__restrict__ rois, int n, const int *n_dev,
                                                              int R, float *__restrict__ wle)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (n_dev) n = *n_dev;
    if (i >= n) return;
    const int W = R * (R + 1), P = R * R;
    const RoiKey a = roi_key(rois, W, P, i);
    float best = 100000000.0f;
    for (int j = 0; j < n; j++) best = fminf(best, nn_term(a, roi_key(rois, W, P, j), j == i));
    wle[(size_t)i * QCS_WLE_FLOATS + QCS_WLE_NEARDIST] = sqrtf(best);
}

// (frame, y)-sorted list: sweep outward through the frame's segment.  Terms from other frames and
// from the ROI itself are >= 1e8 and never beat the scan's initial value 1e8, so only same-frame
// neighbours matter; the sweep stops once the row distance alone exceeds the best distance.
__global__ void __launch_bounds__(256) wle_nn_sorted_kernel(const uint16_t *__restrict__ rois, int n, const int *n_dev,
                                                            int R, float *__restrict__ wle)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (n_dev) n = *n_dev;
    if (i >= n) return;
    const int W = R * (R + 1), P = R * R;
    const RoiKey a = roi_key(rois, W, P, i);
    float best = 100000000.0f;
    for (int j = i + 1; j < n; j++) {
        const RoiKey b = roi_key(rois, W, P, j);
        const int dy = b.y - a.y;
        if (b.f != a.f || (float)(dy * dy) > best) break;
        best = fminf(best, nn_term(a, b, false));
    }
    for (int j = i - 1; j >= 0; j--) {
        const RoiKey b = roi_key(rois, W, P, j);
        const int dy = a.y - b.y;
        if (b.f != a.f || (float)(dy * dy) > best) break;
        best = fminf(best, nn_term(a, b, false));
    }
    wle[(size_t)i * QCS_WLE_FLOATS + QCS_WLE_NEARDIST] = sqrtf(best);
}

// Same sweep over the extraction's (x, y-in-batch, batch) list, which is in raster order: the
// frame of an entry is batch * B + y / H, and two entries share a frame when they share the
// batch and the frame's row range.
// A work-group first copies its entries and NN_HALO entries on either side into LDS; a sweep
// that runs past that window reads the rest from HBM.
constexpr int NN_BLOCK = 256, NN_HALO = 512;
__global__ void __launch_bounds__(NN_BLOCK) wle_nn_xyb_kernel(const int3 *__restrict__ xyb, const int *__restrict__ n_dev,
                                                              int B, int H, float *__restrict__ wle)
{
    __shared__ int3 win[NN_BLOCK + 2 * NN_HALO];
    const int n = *n_dev;
    const int i0 = blockIdx.x * NN_BLOCK;
    if (i0 >= n) return;   // uniform over the work-group
    const int w0 = max(i0 - NN_HALO, 0), w1 = min(i0 + NN_BLOCK + NN_HALO, n);
    for (int j = w0 + (int)threadIdx.x; j < w1; j += NN_BLOCK) win[j - w0] = xyb[j];
    __syncthreads();
    const int i = i0 + threadIdx.x;
    if (i >= n) return;
    const int3 a = win[i - w0];
    const int fy0 = a.y - a.y % H;   // first batch row of a's frame
    float best = 100000000.0f;
    // forward: false once a neighbour ends the sweep
    auto fwd = [&](const int3 q) {
        const int dy = q.y - a.y;
        if (q.z != a.z || q.y - fy0 >= H || (float)(dy * dy) > best) return false;
        const int dx = q.x - a.x;
        best = fminf(best, (float)(dx * dx + dy * dy));
        return true;
    };
    auto bwd = [&](const int3 q) {
        const int dy = a.y - q.y;
        if (q.z != a.z || q.y < fy0 || (float)(dy * dy) > best) return false;
        const int dx = q.x - a.x;
        best = fminf(best, (float)(dx * dx + dy * dy));
        return true;
    };
    int j = i + 1;
    bool go = true;
    for (; go && j < w1; j++) go = fwd(win[j - w0]);
    for (; go && j < n; j++) go = fwd(xyb[j]);
    j = i - 1;
    go = true;
    for (; go && j >= w0; j--) go = bwd(win[j - w0]);
    for (; go && j >= 0; j--) go = bwd(xyb[j]);
    wle[(size_t)i * QCS_WLE_FLOATS + QCS_WLE_NEARDIST] = sqrtf(best);
    (void)B;
}

__global__ void fill_flag_kernel(int *flag) { *flag = 0; }

// 1 in *flag when the list is not sorted by (frame, y)
__global__ void wle_sorted_check_kernel(const uint16_t *__restrict__ rois, int n, int R, int *flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 || i >= n) return;
    const int W = R * (R + 1), P = R * R;
    const RoiKey a = roi_key(rois, W, P, i - 1), b = roi_key(rois, W, P, i);
    if (b.f < a.f || (b.f == a.f && b.y < a.y)) atomicOr(flag, 1);
}

__global__ void __launch_bounds__(256) wle_classify_kernel(const uint16_t *__restrict__ rois, int n, const int *n_dev, int R,
                                                           int loc_type, float *__restrict__ wle)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (n_dev) n = *n_dev;
    if (i >= n) return;
    const int W = R * (R + 1), P = R * R;
    const float th = (loc_type == QCS_LOCTYPE_GS2D) ? ((float)R / 2.1f) / 2.35f : ((float)R / 1.9f) / 2.35f;
    float *o = wle + (size_t)i * QCS_WLE_FLOATS;
    const int det = rois[(size_t)i * W + P + 4];
    const float nn = o[0], sl = o[1], sr = o[2], su = o[3], sd = o[4];
    const float d1 = fabsf(sl - sr) / fminf(sl, sr);
    const float d2 = fabsf(su - sd) / fminf(su, sd);
    const bool contam = (d1 >= 0.40f) || (d2 >= 0.40f);
    const float ms = (((sl + sr) + su) + sd) / 4.0f;
    int mt;
    if (det == QCS_ROITYPE_SINGLE && !contam && nn >= (float)R && ms <= th) mt = QCS_MOLTYPE_MLEFIT;
    else if (det == QCS_ROITYPE_MULTI || ms > th || nn < ((float)R / 2.0f + 1.0f)) mt = QCS_MOLTYPE_MULTIFIT;
    else mt = QCS_MOLTYPE_WLEFIT;
    o[QCS_WLE_FLUOTYPE] = (float)mt;
}

// n ROIs (host count) or, when n_dev is set, *n_dev ROIs with grids sized for n.
// sorted: 2 = extraction list (raster order, (x, y, batch) entries in d_xyb, count in n_dev);
// 1 = list sorted by (frame, y); 0 = unknown order (all-pairs scan).
static int wle_launch_impl(const qcs_params &p, const uint16_t *d_rois, int n, const int *n_dev, float *d_wle, int sorted,
                           hipStream_t s, const int3 *d_xyb = nullptr, int B = 1, int H = 1)
{
    if (n <= 0) return QCS_OK;
    const int R = p.ROISize;
    const dim3 g64((n + 63) / 64), g256((n + 255) / 256);
    switch (R) {
    case 5: hipLaunchKernelGGL(wle_psf_kernel<5>, g64, dim3(64), 0, s, d_rois, n, n_dev, d_wle); break;
    case 7: hipLaunchKernelGGL(wle_psf_kernel<7>, g64, dim3(64), 0, s, d_rois, n, n_dev, d_wle); break;
    case 9: hipLaunchKernelGGL(wle_psf_kernel<9>, g64, dim3(64), 0, s, d_rois, n, n_dev, d_wle); break;
    case 11: hipLaunchKernelGGL(wle_psf_kernel<11>, g64, dim3(64), 0, s, d_rois, n, n_dev, d_wle); break;
    case 13: hipLaunchKernelGGL(wle_psf_kernel<13>, g64, dim3(64), 0, s, d_rois, n, n_dev, d_wle); break;
    case 15: hipLaunchKernelGGL(wle_psf_kernel<15>, g64, dim3(64), 0, s, d_rois, n, n_dev, d_wle); break;
    case 17: hipLaunchKernelGGL(wle_psf_kernel<17>, g64, dim3(64), 0, s, d_rois, n, n_dev, d_wle); break;
    default: return fail(QCS_ERR_ARG, "ROISize %d not supported", R);
    }
    QCS_HIP(hipGetLastError());
    if (sorted == 2)
        hipLaunchKernelGGL(wle_nn_xyb_kernel, dim3((n + NN_BLOCK - 1) / NN_BLOCK), dim3(NN_BLOCK), 0, s, d_xyb, n_dev, B, H, d_wle);
    else if (sorted)
        hipLaunchKernelGGL(wle_nn_sorted_kernel, g256, dim3(256), 0, s, d_rois, n, n_dev, R, d_wle);
    else
        hipLaunchKernelGGL(wle_nn_allpairs_kernel, g256, dim3(256), 0, s, d_rois, n, n_dev, R, d_wle);
    QCS_HIP(hipGetLastError());
    hipLaunchKernelGGL(wle_classify_kernel, g256, dim3(256), 0, s, d_rois, n, n_dev, R, p.LocType, d_wle);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

// extraction output: raster order, count on the device, (x, y, batch) list alongside
int wle_launch_extract(const qcs_params &p, const uint16_t *d_rois, const int3 *d_xyb, int B, int H, const int *d_n, int cap,
                       float *d_wle, hipStream_t s)
{
    return wle_launch_impl(p, d_rois, cap, d_n, d_wle, 2, s, d_xyb, B, H);
}

// arbitrary ROI array: check the order on the device first
int wle_launch(const qcs_params &p, const uint16_t *d_rois, int n, float *d_wle, int *d_flag, hipStream_t s)
{
    if (n <= 0) return QCS_OK;
    hipLaunchKernelGGL(fill_flag_kernel, dim3(1), dim3(1), 0, s, d_flag);
    hipLaunchKernelGGL(wle_sorted_check_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_rois, n, p.ROISize, d_flag);
    QCS_HIP(hipGetLastError());
    int h = 0;
    QCS_HIP(hipMemcpyAsync(&h, d_flag, sizeof(int), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipStreamSynchronize(s));
    return wle_launch_impl(p, d_rois, n, nullptr, d_wle, h == 0, s);
}

}  // namespace qcs

extern "C" int qcs_wle_dev(qcs_ctx *c, const uint16_t *d_rois, int n, float *d_wle, void *stream)
{
    using namespace qcs;
    if (!c || n < 0 || (n > 0 && (!d_rois || !d_wle))) return fail(QCS_ERR_ARG, "invalid argument");
    if (n == 0) return QCS_OK;
    int rc = c->fit_cnt.ensure(64);
    if (rc) return rc;
    return wle_launch(c->p, d_rois, n, d_wle, c->fit_cnt.as<int>(), (hipStream_t)stream);
}
