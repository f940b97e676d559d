// stats.hip -- statistics of a batch of localization rows on gfx950.
//
// Reference (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll):
//   CalcLocResultStatInf       DLL/StatisticsInfo/statisticsInfo.cu:19-91
//   GetStatisticalInf          DLL/StatisticsInfo/statisticsInfo_top.cu:61-128
//   CalcLocalizationDensity2D  DLL/StatisticsInfo/statisticsInfo_top.cu:130-205
//   UpdateStatDat              DLL/StatisticsInfo/statisticsInfo_top.cu:208-225
//   FillerMoleculesInCurFrame, CalcOverlapFluoNum  DLL/StatisticsInfo/statisticsInfo_densityCalc.cu:17-107
//   GetActivationDensity, GetHistogramMeanData     DLL/StatisticsInfo/statisticsInfo_rc.cu:17-74
//   GetFirstFrame, GetLastFrame                    DLL/bfgsMLE_Loc_top/bfgs_top.cu:345-381
//
// The row pass builds the four histograms in LDS per work-group (one global add per non-empty
// bin) and the sums in FP64.  The reference counts neighbours by an all-pairs scan of every
// frame's molecules, one frame per launch with a host synchronisation each; here the molecules
// of all frames are sorted once by (frame, cell of the neighbour radius) and each scans the 3x3
// cells around its own, which visits every pair closer than the radius: the counts are the
// same integers.
#include <limits.h>
#include <stddef.h>
#include <algorithm>
#include <math.h>
#include <string.h>
#include <vector>
#include "qcs_internal.h"

namespace qcs {

namespace {

constexpr int SB = QCS_STAT_BINS, SL = 400;

// device accumulator: histograms, FP64 sums, counters
struct StatAcc {
    int hist[4][SB];
    double sum[4];   // loc prec, peak SNR, PSF width, background
    int valid, first_idx, last_idx, kept, n0, n1, pad[2];
};

// bin of a quotient value / width, clamped to [0, SB - 1] (the reference converts to int with
// saturation and clamps; clamping the float first gives the same bin without an out-of-range
// conversion)
__device__ __forceinline__ int sbin(float q) { return q >= (float)SB ? SB - 1 : (q > 0.0f ? (int)q : 0); }

__global__ void __launch_bounds__(256) stat_rows_kernel(const float *__restrict__ loc, int n, StatAcc *__restrict__ acc)
{
    __shared__ int h[4][SB];
    for (int k = threadIdx.x; k < 4 * SB; k += blockDim.x) (&h[0][0])[k] = 0;
    __syncthreads();
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    int valid = 0, first = INT_MAX, last = -1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float *r = loc + (size_t)i * QCS_REC_FLOATS;
        const int fr = (int)r[QCS_POS_FRME];
        if (fr != 0) {
            first = min(first, i);
            if (i > 0) last = max(last, i);
        }
        const float tp = r[QCS_POS_TPHO], bg = r[QCS_POS_BAKG];
        const float sg = (r[QCS_POS_SIGX] + r[QCS_POS_SIGY]) / 2;
        const float snr = r[QCS_POS_PSNR];
        const float cx = r[QCS_POS_CRBX], cy = r[QCS_POS_CRBY];
        const float lp = sqrtf(cx * cx + cy * cy);
        if (!((tp > 30.0f) && (bg > 2.0f) && (snr >= 3.0f) && (sg >= 0.01f) && (lp > 0.01f) && (lp < 55.0f))) continue;
        atomicAdd(&h[0][sbin(tp / (10000.0f / SL))], 1);
        atomicAdd(&h[1][sbin(lp / (60.0f / SL))], 1);
        atomicAdd(&h[2][sbin(snr / (50.0f / SL))], 1);
        atomicAdd(&h[3][sbin(sg / (5.0f / SL))], 1);
        valid++;
        s[0] += lp;
        s[1] += snr;
        s[2] += sg;
        s[3] += bg;
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int k = 0; k < 4; k++) s[k] += __shfl_down(s[k], off, 64);
        valid += __shfl_down(valid, off, 64);
        first = min(first, __shfl_down(first, off, 64));
        last = max(last, __shfl_down(last, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        for (int k = 0; k < 4; k++)
            if (s[k] != 0.0) atomicAdd(&acc->sum[k], s[k]);
        if (valid) atomicAdd(&acc->valid, valid);
        if (first != INT_MAX) atomicMin(&acc->first_idx, first);
        if (last >= 0) atomicMax(&acc->last_idx, last);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 4 * SB; k += blockDim.x) {
        const int v = (&h[0][0])[k];
        if (v) atomicAdd(&(&acc->hist[0][0])[k], v);
    }
}

// density cell key of every molecule of frames fs..fe at least 5 px inside the image
// (FillerMoleculesInCurFrame); other rows get the key `none`, sorted behind them.  Cells are
// `cell` = radius * (1 + 2^-10) wide: two molecules whose float distance test passes are then at
// most one cell apart in x and in y even after the rounding of the quotients.
__global__ void __launch_bounds__(256) stat_keys_kernel(const float *__restrict__ loc, int n, int fs, int fe, int W, int H,
                                                        float cell, int ncx, int ncy, unsigned long long none,
                                                        unsigned long long *__restrict__ keys, int *__restrict__ vals,
                                                        StatAcc *__restrict__ acc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool kept = false;
    if (i < n) {
        const float *r = loc + (size_t)i * QCS_REC_FLOATS;
        const float X = r[QCS_POS_XPOS], Y = r[QCS_POS_YPOS];
        const int x = (int)X, y = (int)Y, cf = (int)r[QCS_POS_FRME];
        unsigned long long key = none;
        if (x >= 5 && x < W - 5 && y >= 5 && y < H - 5 && cf >= fs && cf <= fe) {
            const int cx = (int)(X / cell), cy = (int)(Y / cell);
            key = ((unsigned long long)(cf - fs) * (unsigned long long)ncy + (unsigned long long)cy) * (unsigned long long)ncx +
                  (unsigned long long)cx;
            kept = true;
        }
        keys[i] = key;
        vals[i] = i;
    }
    const unsigned long long m = __ballot(kept);
    if (m && (threadIdx.x & 63) == 0) atomicAdd(&acc->kept, __popcll(m));
}

__device__ __forceinline__ int lower_bound_key(const unsigned long long *__restrict__ keys, int m, unsigned long long k)
{
    int lo = 0, hi = m;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// neighbours within the radius in the same frame (CalcOverlapFluoNum): molecules with 0 and 1
__global__ void __launch_bounds__(256) stat_neigh_kernel(const float *__restrict__ loc, const unsigned long long *__restrict__ keys,
                                                         const int *__restrict__ vals, int m, int ncx, int ncy, float r2,
                                                         StatAcc *__restrict__ acc)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    int c0 = 0, c1 = 0;
    if (p < m) {
        const unsigned long long key = keys[p];
        const int row = vals[p];
        const int cx = (int)(key % (unsigned long long)ncx);
        const unsigned long long fy = key / (unsigned long long)ncx;
        const int cy = (int)(fy % (unsigned long long)ncy);
        const unsigned long long fbase = (fy / (unsigned long long)ncy) * (unsigned long long)ncy;
        const float X = loc[(size_t)row * QCS_REC_FLOATS + QCS_POS_XPOS], Y = loc[(size_t)row * QCS_REC_FLOATS + QCS_POS_YPOS];
        int cnt = 0;
        for (int dy = -1; dy <= 1; dy++) {
            const int yy = cy + dy;
            if (yy < 0 || yy >= ncy) continue;
            for (int dx = -1; dx <= 1; dx++) {
                const int xx = cx + dx;
                if (xx < 0 || xx >= ncx) continue;
                const unsigned long long k2 = (fbase + (unsigned long long)yy) * (unsigned long long)ncx + (unsigned long long)xx;
                for (int q = lower_bound_key(keys, m, k2); q < m && keys[q] == k2; q++) {
                    const int o = vals[q];
                    if (o == row) continue;
                    const float ox = loc[(size_t)o * QCS_REC_FLOATS + QCS_POS_XPOS] - X;
                    const float oy = loc[(size_t)o * QCS_REC_FLOATS + QCS_POS_YPOS] - Y;
                    cnt += (ox * ox + oy * oy) <= r2;
                }
            }
        }
        c0 = cnt == 0;
        c1 = cnt == 1;
    }
    for (int off = 32; off > 0; off >>= 1) {
        c0 += __shfl_down(c0, off, 64);
        c1 += __shfl_down(c1, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (c0) atomicAdd(&acc->n0, c0);
        if (c1) atomicAdd(&acc->n1, c1);
    }
}

__global__ void stat_init_kernel(StatAcc *acc)
{
    // one writer per word: first_idx starts at INT_MAX, last_idx at -1, everything else at 0
    int *p = reinterpret_cast<int *>(acc);
    const int nw = (int)(sizeof(StatAcc) / sizeof(int));
    const int wf = (int)(offsetof(StatAcc, first_idx) / sizeof(int)), wl = (int)(offsetof(StatAcc, last_idx) / sizeof(int));
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nw; k += gridDim.x * blockDim.x)
        p[k] = k == wf ? INT_MAX : (k == wl ? -1 : 0);
}

// GetHistogramMeanData (statisticsInfo_rc.cu:44-74), float arithmetic in bin order
float hist_mean(const int *h, int len, float pct)
{
    int mx = h[0];
    for (int i = 0; i < len; i++) mx = h[i] > mx ? h[i] : mx;
    pct = pct * (float)mx;
    float ws = 0.0f, os = 0.0f;
    for (int i = 0; i < len; i++) {
        const float v = (float)h[i];
        if (v >= pct) {
            ws += v * ((float)i + 0.5f);
            os += v;
        }
    }
    if (os == 0) os = 1;
    return ws / os;
}

// GetActivationDensity (statisticsInfo_rc.cu:17-41)
float activation_density(float r0, float r1, float radius_um)
{
    const float PI = 3.14159f;
    if (r0 <= 0.001f) r0 = 0.001f;
    if (r1 <= 0.001f) r1 = 0.001f;
    if (r0 > 1) r0 = 1;
    if (r1 > 1) r1 = 1;
    const float l0 = logf(r0);
    const float l1 = logf(r1 / PI / radius_um / radius_um);
    const float d0 = -l0 / PI / radius_um / radius_um;
    const float d1 = -l1 / PI / radius_um / radius_um;
    return r0 * d0 + (1.0f - r0) * d1;
}

}  // namespace

}  // namespace qcs

using namespace qcs;

extern "C" int qcs_loc_stats_dev(qcs_ctx *ctx, const float *d_loc, int n, qcs_loc_stats *out, void *stream)
{
    if (!ctx || !out || n < 0 || (n > 0 && !d_loc)) return fail(QCS_ERR_ARG, "invalid argument");
    memset(out, 0, sizeof(*out));
    if (n == 0) return QCS_OK;
    hipStream_t s = (hipStream_t)stream;
    const qcs_params &p = ctx->p;
    // scratch: accumulator, keys and values (double-buffered for the sort), sort workspace
    const size_t temp = sizeof(int) * radix_sort_tmp_ints(n);
    const size_t acc_b = (sizeof(StatAcc) + 255) & ~(size_t)255, keys_b = ((size_t)n * 8 + 255) & ~(size_t)255,
                 vals_b = ((size_t)n * 4 + 255) & ~(size_t)255;
    int rc;
    if ((rc = ctx->st_buf.ensure(acc_b + 2 * keys_b + 2 * vals_b + temp + 256))) return rc;
    char *b = ctx->st_buf.as<char>();
    StatAcc *acc = reinterpret_cast<StatAcc *>(b);
    unsigned long long *k_in = reinterpret_cast<unsigned long long *>(b + acc_b), *k_out = reinterpret_cast<unsigned long long *>(b + acc_b + keys_b);
    int *v_in = reinterpret_cast<int *>(b + acc_b + 2 * keys_b), *v_out = reinterpret_cast<int *>(b + acc_b + 2 * keys_b + vals_b);
    void *tmp = b + acc_b + 2 * keys_b + 2 * vals_b;
    hipLaunchKernelGGL(stat_init_kernel, dim3(8), dim3(256), 0, s, acc);
    const int blocks = (int)std::min<size_t>(((size_t)n + 255) / 256, 1024);
    hipLaunchKernelGGL(stat_rows_kernel, dim3(blocks), dim3(256), 0, s, d_loc, n, acc);
    QCS_HIP(hipGetLastError());
    std::vector<char> hb(sizeof(StatAcc));
    StatAcc *h = reinterpret_cast<StatAcc *>(hb.data());
    QCS_HIP(hipMemcpyAsync(h, acc, sizeof(StatAcc), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipStreamSynchronize(s));
    memcpy(out->hist_total_photon, h->hist[0], sizeof(out->hist_total_photon));
    memcpy(out->hist_loc_prec, h->hist[1], sizeof(out->hist_loc_prec));
    memcpy(out->hist_peak_snr, h->hist[2], sizeof(out->hist_peak_snr));
    memcpy(out->hist_psf_width, h->hist[3], sizeof(out->hist_psf_width));
    out->valid = h->valid;
    out->sum_loc_prec = h->sum[0];
    out->sum_peak_snr = h->sum[1];
    out->sum_psf_width = h->sum[2];
    out->sum_background = h->sum[3];
    const double vn = h->valid > 0 ? (double)h->valid : 1.0;
    out->mean_total_photon = (10000.0f / SL) * hist_mean(out->hist_total_photon, SB, 0.1f);
    out->mean_loc_prec = (float)(out->sum_loc_prec / vn);
    out->mean_peak_snr = (float)(out->sum_peak_snr / vn);
    out->mean_background = (float)(out->sum_background / vn);
    out->mean_psf_width = (float)(out->sum_psf_width / vn);
    out->mean_psf_width_ctl = out->mean_psf_width * (-out->mean_peak_snr);

    // frame range (GetFirstFrame / GetLastFrame), both ends dropped when more than 3 frames
    int fs = 0, fe = 0;
    float fv;
    if (h->first_idx != INT_MAX) {
        QCS_HIP(hipMemcpy(&fv, d_loc + (size_t)h->first_idx * QCS_REC_FLOATS + QCS_POS_FRME, sizeof(float), hipMemcpyDeviceToHost));
        fs = (int)fv;
    }
    if (h->last_idx >= 0) {
        QCS_HIP(hipMemcpy(&fv, d_loc + (size_t)h->last_idx * QCS_REC_FLOATS + QCS_POS_FRME, sizeof(float), hipMemcpyDeviceToHost));
        fe = (int)fv;
    }
    if (fe - fs + 1 > 3) {
        fs++;
        fe--;
    }
    out->density_first_frame = fs;
    out->density_last_frame = fe;
    if (fe < fs) return QCS_OK;
    const float rad = (p.LocType == QCS_LOCTYPE_GS2D) ? 10.0f : 20.0f;
    const float cell = rad * 1.0009765625f;
    const int ncx = (int)((float)p.ImageWidth / cell) + 2, ncy = (int)((float)p.ImageHigh / cell) + 2;
    const unsigned long long none = (unsigned long long)(fe - fs + 1) * (unsigned long long)ncy * (unsigned long long)ncx;
    int bits = 1;
    while (bits < 64 && (none >> bits) != 0ull) bits++;
    const dim3 g((n + 255) / 256), blk(256);
    hipLaunchKernelGGL(stat_keys_kernel, g, blk, 0, s, d_loc, n, fs, fe, p.ImageWidth, p.ImageHigh, cell, ncx, ncy, none, k_in, v_in, acc);
    QCS_HIP(hipGetLastError());
    unsigned long long *k_sorted;
    int *v_sorted;
    if ((rc = radix_sort_pairs_u64(k_in, v_in, k_out, v_out, n, bits, static_cast<int *>(tmp), s, &k_sorted, &v_sorted))) return rc;
    int kept = 0;
    QCS_HIP(hipMemcpyAsync(&kept, &acc->kept, sizeof(int), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipStreamSynchronize(s));
    if (kept > 0) {
        hipLaunchKernelGGL(stat_neigh_kernel, dim3((kept + 255) / 256), blk, 0, s, d_loc, k_sorted, v_sorted, kept, ncx, ncy, rad * rad,
                           acc);
        QCS_HIP(hipGetLastError());
    }
    int cnt[2] = {0, 0};
    QCS_HIP(hipMemcpyAsync(cnt, &acc->n0, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipStreamSynchronize(s));
    out->density_total = kept;
    out->density_n0 = cnt[0];
    out->density_n1 = cnt[1];
    if (kept > 50) {
        const float r0 = (float)cnt[0] / (float)kept, r1 = (float)cnt[1] / (float)kept;
        float d = activation_density(r0, r1, rad * p.PixelSize / 1000);
        d = p.MultiEmitterFitEn ? d * d * 0.5098f + d * 0.6364f : d * d * 0.8082f + d * 0.519f;
        out->loc_density_2d = d;
    }
    return QCS_OK;
}
