// extract.hip -- ROI extraction from uint16 sCMOS frames on gfx950.
//
// Reference (SE = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll/SubregionExtraction):
//   FilterInit                    SE/LDROIExtraction_ImageFiltering.cu:22-81
//   ImageFiltering pass order     SE/LDROIExtraction_ImageFiltering.cu:85-121
//   gpuRemoveBackgroundIntensity  SE/LDROIExtraction_ImageFiltering.cu:125-282
//   gpuImageConv_Seperable_H / V  SE/LDROIExtraction_ImageFiltering.cu:286-470
//   ImageVarianceCalc             SE/LDROIExtraction.cu:140-177, ..._ImageFiltering.cu:503-544
//   gpuROIFindingLD               SE/LDROIExtraction_ROIExtraction.cu:25-177
//   gpuDetectedROIClasify         SE/LDROIExtraction_ROIExtraction.cu:181-253
//   gpuSubregionExtractionLD      SE/LDROIExtraction_ROIExtraction.cu:257-316
//   ExtractMolecules / ROIExtraction SE/LDROIExtraction.cu:30-138
//
// Pipeline for one call (any number of frames, batches of BatchedImageNum frames; no host
// synchronisation until the final count read):
//   1. var_kernel     exact integer sums of the background-removed variance samples per batch
//   2. detect_kernel  one 64 x 16 output tile per work-group: raw tile + halo -> LDS, background
//                     removal, 5-tap horizontal then vertical smoothing (uint16 truncation after
//                     each pass, as the reference stores them), 3x3 local-maximum test against
//                     the batch's thresholds.  Candidates and marks leave as one 64-bit ballot
//                     mask per row segment; per-row candidate counts via integer atomics.
//   3. scan           exclusive prefix sum of the row counts -> raster-order offsets
//   4. batch_kernel   per-batch found / kept (capacity) / output offsets
//   5. emit_kernel    candidate (x, y) list in raster order
//   6. cut_kernel     detector classification (multi mode), raw R x R cut, header
//   7. WLE            PSF width, nearest neighbour (sorted sweep), type
#include <string.h>
#include "qcs_internal.h"

namespace qcs {

int wle_launch_devn(const qcs_params &p, const uint16_t *d_rois, const int *d_n, int cap, float *d_wle, hipStream_t s);
void filter_taps_host(int R, int loc_type, int multi, float taps[5]);

// FilterInit (SE/LDROIExtraction_ImageFiltering.cu:22-81): host FP32 without contraction; the
// exponential is evaluated in double and rounded once.
void filter_taps_host(int R, int loc_type, int multi, float taps[5])
{
    volatile float s7 = 0.5f, s11 = 1.0f;
    float ss = s7 + ((float)(R - 7) * (s11 - s7)) / (float)(11 - 7);
    if (ss < 0.5f) ss = 0.5f;
    if (loc_type == QCS_LOCTYPE_AS3D) ss = ss * 1.5f;
    else if (multi == 0 && loc_type == QCS_LOCTYPE_GS2D) ss = ss * 1.2f;
    ss = -1.0f / ((2.0f * ss) * ss);
    const float ctr = 2.0f;
    for (int c = 0; c < 5; c++) {
        volatile float d = ((float)c - ctr) * ((float)c - ctr);
        volatile float arg = d * ss;
        taps[c] = (float)exp((double)arg);
    }
    float sum = 0.0f;
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            volatile float pr = taps[i] * taps[j];
            sum = sum + pr;
        }
    sum = sqrtf(1.0f / sum);
    for (int c = 0; c < 5; c++) taps[c] = taps[c] * sum;
}

struct ExtractConst {
    int W, H;              // frame size
    int B;                 // frames per batch
    int nframes;           // frames in this call
    int nbatch;
    int R;
    int loc_type, multi;
    int smooth;            // 0: detection on the background-removed image (2D multi mode)
    int batch_cap;         // candidates kept per batch
    int cap;               // output capacity
    int start_frame;
    int nseg;              // 64-pixel segments per row
    float taps[5];
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// rows of batch b and its first frame
__device__ __forceinline__ int batch_rows(const ExtractConst &c, int b)
{
    const int f0 = b * c.B;
    const int nb = (c.nframes - f0) < c.B ? (c.nframes - f0) : c.B;
    return nb * c.H;
}

// sum over 8 positions of the minimum of four border samples of the 9 x 9 window (exact in
// integers; the reference's float sum of uint16 minima is exact too), from global memory
__device__ __forceinline__ int bkg_sum_global(const uint16_t *img, int W, int BH, int x, int y)
{
    const int yt = clampi(y - 4, 0, BH - 1), yb = clampi(y + 4, 0, BH - 1);
    const int xl = clampi(x - 4, 0, W - 1), xr = clampi(x + 4, 0, W - 1);
    int s = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int l1 = img[(size_t)yt * W + clampi(x - 4 + i, 0, W - 1)];
        const int l2 = img[(size_t)clampi(y - 4 + i, 0, BH - 1) * W + xr];
        const int l3 = img[(size_t)yb * W + clampi(x - 3 + (7 - i), 0, W - 1)];
        const int l4 = img[(size_t)clampi(y - 3 + (7 - i), 0, BH - 1) * W + xl];
        const int a = l1 < l2 ? l1 : l2, b = l3 < l4 ? l3 : l4;
        s += a < b ? a : b;
    }
    return s;
}

// background-removed value (uint16 truncation) of in-image pixel (x, y), from global memory
__device__ __forceinline__ int bkg_removed_global(const uint16_t *img, int W, int BH, int x, int y)
{
    const int g0 = x & ~31;
    const int xl = (x - 1 < g0) ? g0 : x - 1, xr = (x + 1 > g0 + 31) ? g0 + 31 : x + 1;
    const float b0 = (float)bkg_sum_global(img, W, BH, xl, y) / 8.0f;
    const float b1 = (float)bkg_sum_global(img, W, BH, x, y) / 8.0f;
    const float b2 = (float)bkg_sum_global(img, W, BH, xr, y) / 8.0f;
    const float mb = ((b0 + b1) + b2) / 3.0f;
    float v = (float)img[(size_t)y * W + x] - mb;
    v = fmaxf(v, 0.0f);
    return (int)v;
}

// ---- 1. variance sums per batch -------------------------------------------------------------
__global__ void __launch_bounds__(256) var_kernel(const uint16_t *__restrict__ frames, ExtractConst c,
                                                  unsigned long long *__restrict__ sums)
{
    const int b = blockIdx.y;
    const int BH = batch_rows(c, b);
    const uint16_t *img = frames + (size_t)b * c.B * c.H * c.W;
    const int cx = c.W / 2, cy = BH / 2;
    int sw = c.W / 2 - 5, sh = BH / 2 - 5;
    sw = sw > 256 ? 256 : sw;
    sh = sh > 256 ? 256 : sh;
    sw = sw < 1 ? 1 : sw;
    sh = sh < 1 ? 1 : sh;
    const int S = sw < sh ? sw : sh;
    unsigned long long s1 = 0, s2 = 0;
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < S * S; g += gridDim.x * blockDim.x) {
        const int X = cx + g % S, Y = cy + g / S;
        int d1 = bkg_removed_global(img, c.W, BH, X, Y);
        const int d2 = bkg_removed_global(img, c.W, BH, X - S, Y);
        int d3 = bkg_removed_global(img, c.W, BH, X, Y - S);
        const int d4 = bkg_removed_global(img, c.W, BH, X - S, Y - S);
        d1 = d1 < d2 ? d1 : d2;
        d3 = d3 < d4 ? d3 : d4;
        d1 = d1 < d3 ? d1 : d3;
        s1 += (unsigned long long)d1;
        s2 += (unsigned long long)d1 * (unsigned long long)d1;
    }
    // wave reduction then one atomic per wave (integer: order-independent)
    for (int off = 32; off > 0; off >>= 1) {
        s1 += __shfl_down(s1, off, 64);
        s2 += __shfl_down(s2, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&sums[2 * b + 0], s1);
        atomicAdd(&sums[2 * b + 1], s2);
    }
}

__device__ __forceinline__ void batch_thresholds(const ExtractConst &c, const unsigned long long *sums, int b, int BH,
                                                 float &th, float &th1)
{
    int sw = c.W / 2 - 5, sh = BH / 2 - 5;
    sw = sw > 256 ? 256 : sw;
    sh = sh > 256 ? 256 : sh;
    sw = sw < 1 ? 1 : sw;
    sh = sh < 1 ? 1 : sh;
    const int S = sw < sh ? sw : sh;
    const double n = (double)S * (double)S;
    const double m1 = (double)sums[2 * b] / n, m2 = (double)sums[2 * b + 1] / n;
    float var = (float)(m2 - m1 * m1);
    var = (var < 100.0f) ? 100.0f : var;
    const float sd = sqrtf(var);
    if (c.loc_type == QCS_LOCTYPE_AS3D || (c.multi == 0 && c.loc_type == QCS_LOCTYPE_GS2D)) {
        th = 8.5f * sd;
        th1 = 3.5f * sd;
    } else {
        th = 9.5f * sd;
        th1 = 3.8f * sd;
    }
}

// ---- 2. fused filter + detect ----------------------------------------------------------------
constexpr int TW = 64, TH = 16;                    // output tile
constexpr int RAW_W = TW + 16, RAW_H = TH + 14;   // raw halo: cols x0-8.., rows y0-7..
constexpr int BK_W = TW + 8, BK_H = TH + 6;       // background sums: cols x0-4.., rows y0-3..
constexpr int T_W = TW + 6;                       // background-removed: cols x0-3.., rows y0-3..
constexpr int HC_W = TW + 2;                      // H-smoothed: cols x0-1.., rows y0-3..
constexpr int S_H = TH + 2;                       // V-smoothed: cols x0-1.., rows y0-1..

__global__ void __launch_bounds__(256) detect_kernel(const uint16_t *__restrict__ frames, ExtractConst c,
                                                     const unsigned long long *__restrict__ var_sums,
                                                     unsigned long long *__restrict__ cand_mask,
                                                     unsigned long long *__restrict__ mark_mask,
                                                     int *__restrict__ row_count)
{
    __shared__ uint16_t raw[RAW_H][RAW_W];
    __shared__ int bk[BK_H][BK_W];
    __shared__ uint16_t tv[BK_H][T_W];
    __shared__ uint16_t hc[BK_H][HC_W];
    __shared__ uint16_t sv[S_H][HC_W];

    const int b = blockIdx.z;
    const int BH = batch_rows(c, b);
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    if (y0 >= BH) return;
    const int W = c.W;
    const uint16_t *img = frames + (size_t)b * c.B * c.H * W;
    const int tid = threadIdx.x;

    // raw tile with clamped coordinates: raw[i][j] = img(clamp(y0-7+i), clamp(x0-8+j))
    for (int k = tid; k < RAW_H * RAW_W; k += 256) {
        const int i = k / RAW_W, j = k % RAW_W;
        raw[i][j] = img[(size_t)clampi(y0 - 7 + i, 0, BH - 1) * W + clampi(x0 - 8 + j, 0, W - 1)];
    }
    __syncthreads();
    // background sums at rows y0-3+i, cols x0-4+j (columns past the right edge keep the
    // reference's clamped loads; rows outside the batch are never read)
    for (int k = tid; k < BK_H * BK_W; k += 256) {
        const int i = k / BK_W, j = k % BK_W;
        const int ri = i + 4, cj = j + 4;   // position in raw[]
        int s = 0;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int l1 = raw[ri - 4][cj - 4 + t];
            const int l2 = raw[ri - 4 + t][cj + 4];
            const int l3 = raw[ri + 4][cj - 3 + (7 - t)];
            const int l4 = raw[ri - 3 + (7 - t)][cj - 4];
            const int a = l1 < l2 ? l1 : l2, bb = l3 < l4 ? l3 : l4;
            s += a < bb ? a : bb;
        }
        bk[i][j] = s;
    }
    __syncthreads();
    // background-removed values at in-image rows y0-3+i, cols x0-3+j
    for (int k = tid; k < BK_H * T_W; k += 256) {
        const int i = k / T_W, j = k % T_W;
        const int y = y0 - 3 + i, x = x0 - 3 + j;
        if (y < 0 || y >= BH || x < 0 || x >= W) continue;
        const int g0 = x & ~31;
        const int xl = (x - 1 < g0) ? g0 : x - 1, xr = (x + 1 > g0 + 31) ? g0 + 31 : x + 1;
        const float b0 = (float)bk[i][xl - (x0 - 4)] / 8.0f;
        const float b1 = (float)bk[i][x - (x0 - 4)] / 8.0f;
        const float b2 = (float)bk[i][xr - (x0 - 4)] / 8.0f;
        const float mb = ((b0 + b1) + b2) / 3.0f;
        float v = (float)raw[i + 4][j + 5] - mb;
        v = fmaxf(v, 0.0f);
        tv[i][j] = (uint16_t)v;
    }
    __syncthreads();
    if (c.smooth) {
        // horizontal pass at rows y0-3+i, cols x0-1+j (clamp-to-edge reads of the removed image)
        for (int k = tid; k < BK_H * HC_W; k += 256) {
            const int i = k / HC_W, j = k % HC_W;
            const int y = y0 - 3 + i, x = x0 - 1 + j;
            if (y < 0 || y >= BH || x < 0 || x >= W) continue;
            float acc = 0.0f;
#pragma unroll
            for (int t = 0; t < 5; t++)
                acc = __builtin_fmaf((float)tv[i][clampi(x - 2 + t, 0, W - 1) - (x0 - 3)], c.taps[t], acc);
            acc = acc > 65535.0f ? 65535.0f : acc;
            hc[i][j] = (uint16_t)acc;
        }
        __syncthreads();
        // vertical pass at rows y0-1+i, cols x0-1+j
        for (int k = tid; k < S_H * HC_W; k += 256) {
            const int i = k / HC_W, j = k % HC_W;
            const int y = y0 - 1 + i, x = x0 - 1 + j;
            if (y < 0 || y >= BH || x < 0 || x >= W) continue;
            float acc = 0.0f;
#pragma unroll
            for (int t = 0; t < 5; t++)
                acc = __builtin_fmaf((float)hc[clampi(y - 2 + t, 0, BH - 1) - (y0 - 3)][j], c.taps[t], acc);
            acc = acc > 65535.0f ? 65535.0f : acc;
            sv[i][j] = (uint16_t)acc;
        }
    } else {
        for (int k = tid; k < S_H * HC_W; k += 256) {
            const int i = k / HC_W, j = k % HC_W;
            const int y = y0 - 1 + i, x = x0 - 1 + j;
            if (y < 0 || y >= BH || x < 0 || x >= W) continue;
            sv[i][j] = tv[i + 2][j + 2];
        }
    }
    __syncthreads();

    // detection: wave w handles rows y0 + 4w .. +3, lane = column
    float th, th1;
    batch_thresholds(c, var_sums, b, BH, th, th1);
    const int lane = tid & 63, wv = tid >> 6;
    const int half = c.R / 2;
    const int x = x0 + lane;
    for (int rr = 0; rr < 4; rr++) {
        const int y = y0 + wv * 4 + rr;
        if (y >= BH) break;   // uniform across the wave
        const int ry = y % c.H;
        bool cand = false, mark = false;
        if (x < W && x >= half && ry >= half && x < W - half && ry < c.H - half) {
            const int si = y - (y0 - 1), sj = x - (x0 - 1);
            const int cv = sv[si][sj];
            if ((float)cv >= th) {
                const int p00 = sv[si - 1][sj - 1], p01 = sv[si - 1][sj], p02 = sv[si - 1][sj + 1];
                const int p10 = sv[si][sj - 1], p12 = sv[si][sj + 1];
                const int p20 = sv[si + 1][sj - 1], p21 = sv[si + 1][sj], p22 = sv[si + 1][sj + 1];
                const int s8 = ((float)p00 > th1) + ((float)p01 > th1) + ((float)p02 > th1) + ((float)p10 > th1) +
                               ((float)p12 > th1) + ((float)p20 > th1) + ((float)p21 > th1) + ((float)p22 > th1);
                const int sin = (cv > p01) + (cv >= p12) + (cv >= p21) + (cv > p10);
                const int sout = (cv > p00) + (cv > p02) + (cv > p20) + (cv > p22);
                mark = (s8 >= 7 && sin == 4 && sout >= 3);
                cand = mark && (sout == 4);
            }
        }
        const unsigned long long cm = __ballot(cand);
        const unsigned long long mm = __ballot(mark);
        if (lane == 0) {
            const size_t grow = (size_t)b * c.B * c.H + y;   // stacked row index
            cand_mask[grow * c.nseg + blockIdx.x] = cm;
            mark_mask[grow * c.nseg + blockIdx.x] = mm;
            if (cm) atomicAdd(&row_count[grow], __popcll(cm));
        }
    }
}

// ---- 3. exclusive scan of int counts (3-pass) -----------------------------------------------
constexpr int SCAN_BLOCK = 256, SCAN_ITEMS = 16, SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__device__ __forceinline__ int block_exclusive_scan(int v, int *sh)
{
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
        const int t = (tid >= off) ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += t;
        __syncthreads();
    }
    const int incl = sh[tid];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(SCAN_BLOCK) scan_tiles_kernel(const int *__restrict__ in, int n, int *__restrict__ out,
                                                                int *__restrict__ tile_sums)
{
    __shared__ int sh[SCAN_BLOCK];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS], s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int run = block_exclusive_scan(s, sh);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == SCAN_BLOCK - 1) tile_sums[blockIdx.x] = run;
}

// single work-group scan of the tile sums (ntiles <= SCAN_TILE), total at tile_sums[ntiles]
__global__ void __launch_bounds__(SCAN_BLOCK) scan_sums_kernel(int *__restrict__ tile_sums, int ntiles)
{
    __shared__ int sh[SCAN_BLOCK];
    const int base = threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS], s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = (base + k < ntiles) ? tile_sums[base + k] : 0;
        s += v[k];
    }
    int run = block_exclusive_scan(s, sh);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < ntiles) tile_sums[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == SCAN_BLOCK - 1) tile_sums[ntiles] = run;
}

__global__ void __launch_bounds__(256) scan_add_kernel(int *__restrict__ out, int n, const int *__restrict__ tile_sums)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += tile_sums[i / SCAN_TILE];
}

// ---- 4. per-batch bookkeeping ----------------------------------------------------------------
// info layout: [0] kept total, [1] found total, then per batch b: found, kept, out offset
__global__ void batch_kernel(ExtractConst c, const int *__restrict__ row_off, const int *__restrict__ tile_sums,
                             int nrows, int ntiles, int *__restrict__ info)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int total = tile_sums[ntiles];
    int out = 0, found_all = 0;
    for (int b = 0; b < c.nbatch; b++) {
        const int r0 = b * c.B * c.H;
        const int r1 = r0 + batch_rows(c, b);
        const int o0 = row_off[r0];
        const int o1 = (r1 < nrows) ? row_off[r1] : total;
        const int found = o1 - o0;
        int kept = found < c.batch_cap ? found : c.batch_cap;
        const int room = c.cap - out;
        kept = kept < room ? kept : (room > 0 ? room : 0);
        info[2 + 3 * b + 0] = found;
        info[2 + 3 * b + 1] = kept;
        info[2 + 3 * b + 2] = out;
        out += kept;
        found_all += found;
    }
    info[0] = out;
    info[1] = found_all;
}

// ---- 5. candidate list in raster order -------------------------------------------------------
__global__ void __launch_bounds__(256) emit_kernel(ExtractConst c, const unsigned long long *__restrict__ cand_mask,
                                                   const int *__restrict__ row_off, const int *__restrict__ row_count,
                                                   int nrows, const int *__restrict__ info, int3 *__restrict__ xyb)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nrows || row_count[row] == 0) return;
    const int b = row / (c.B * c.H);
    const int r0 = b * c.B * c.H;
    const int kept = info[2 + 3 * b + 1], out0 = info[2 + 3 * b + 2];
    int local = row_off[row] - row_off[r0];
    const int y = row - r0;   // batch-relative row
    for (int sgi = 0; sgi < c.nseg; sgi++) {
        unsigned long long m = cand_mask[(size_t)row * c.nseg + sgi];
        while (m) {
            const int bit = __ffsll((long long)m) - 1;
            m &= m - 1;
            if (local < kept) xyb[out0 + local] = make_int3(sgi * 64 + bit, y, b);
            local++;
        }
    }
}

// ---- 6. classification + cut -----------------------------------------------------------------
__device__ __forceinline__ int mark_count(const unsigned long long *mark, int nseg, size_t row, int xa, int xb)
{
    // number of marks in [xa, xb] of one stacked row
    int cnt = 0;
    for (int sgi = xa >> 6; sgi <= (xb >> 6); sgi++) {
        unsigned long long m = mark[row * nseg + sgi];
        const int lo = sgi * 64;
        if (xa > lo) m &= ~0ULL << (xa - lo);
        if (xb < lo + 63) m &= ~0ULL >> (63 - (xb - lo));
        cnt += __popcll(m);
    }
    return cnt;
}

__global__ void __launch_bounds__(256) cut_kernel(const uint16_t *__restrict__ frames, ExtractConst c,
                                                  const unsigned long long *__restrict__ mark_mask,
                                                  const int *__restrict__ info, const int3 *__restrict__ xyb,
                                                  uint16_t *__restrict__ rois)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= info[0]) return;
    const int R = c.R, half = R / 2, WR = R * (R + 1);
    const int3 e = xyb[i];
    const int x = e.x, y = e.y, b = e.z;
    const size_t rbase = (size_t)b * c.B * c.H;   // stacked row of the batch's first row
    int ft = QCS_ROITYPE_SINGLE;
    if (c.multi) {
        int all = 0, inner = 0;
        for (int r = 0; r < R; r++) all += mark_count(mark_mask, c.nseg, rbase + y - half + r, x - half, x + half);
        for (int r = 0; r < 3; r++) inner += mark_count(mark_mask, c.nseg, rbase + y - 1 + r, x - 1, x + 1);
        ft = (all - inner > 0) ? QCS_ROITYPE_MULTI : QCS_ROITYPE_SINGLE;
    }
    const uint16_t *img = frames + rbase * c.W;
    uint16_t *o = rois + (size_t)i * WR;
    for (int r = 0; r < R; r++)
        for (int cc = 0; cc < R; cc++) o[r * R + cc] = img[(size_t)(y - half + r) * c.W + (x - half + cc)];
    const int fr = c.start_frame + b * c.B + y / c.H;
    o[R * R + 0] = (uint16_t)x;
    o[R * R + 1] = (uint16_t)(y % c.H);
    o[R * R + 2] = (uint16_t)(fr % 65536);
    o[R * R + 3] = (uint16_t)(fr / 65536);
    o[R * R + 4] = (uint16_t)ft;
    for (int k = 5; k < R; k++) o[R * R + k] = 0;
}

__global__ void fill_kernel(int *__restrict__ p, size_t n, int v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

void fill_async(int *p, size_t n, int v, hipStream_t s)
{
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

int scan_tile_count(size_t n) { return (int)((n + SCAN_TILE - 1) / SCAN_TILE); }

int scan_exclusive(const int *in, int n, int *out, int *tile_sums, hipStream_t s)
{
    if (n <= 0) return QCS_OK;
    const int ntiles = scan_tile_count((size_t)n);
    if (ntiles > SCAN_TILE) return fail(QCS_ERR_ARG, "scan of %d items exceeds the supported size", n);
    hipLaunchKernelGGL(scan_tiles_kernel, dim3(ntiles), dim3(SCAN_BLOCK), 0, s, in, n, out, tile_sums);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(SCAN_BLOCK), 0, s, tile_sums, ntiles);
    hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, n, tile_sums);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

__global__ void __launch_bounds__(256) scatter_kernel(const int *__restrict__ flags, const int *__restrict__ offs, int n,
                                                      const int *__restrict__ src, const int *__restrict__ base_dev,
                                                      int base_const, int *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || !flags[j]) return;
    const int base = base_dev ? *base_dev : base_const;
    out[base + offs[j]] = src ? src[j] : j;
}

int compact_scatter(const int *flags, const int *offs, int n, const int *src, const int *base_dev, int base_const, int *out,
                    hipStream_t s)
{
    if (n <= 0) return QCS_OK;
    hipLaunchKernelGGL(scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, s, flags, offs, n, src, base_dev, base_const, out);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

}  // namespace qcs

using namespace qcs;

extern "C" int qcs_extract_dev(qcs_ctx *ctx, const uint16_t *d_frames, int nframes, int start_frame, uint16_t *d_rois,
                               float *d_wle, int cap, int *n_out, int *n_found, void *stream)
{
    if (n_out) *n_out = 0;
    if (n_found) *n_found = 0;
    if (!ctx || nframes < 0 || cap < 0 || (nframes > 0 && !d_frames) || (cap > 0 && (!d_rois || !d_wle)))
        return fail(QCS_ERR_ARG, "invalid argument");
    const qcs_params &p = ctx->p;
    if (p.ImageWidth <= 0 || p.ImageHigh <= 0) return fail(QCS_ERR_ARG, "ImageWidth/ImageHigh not set");
    const int R = p.ROISize;
    if (p.ImageWidth < R || p.ImageHigh < R) return fail(QCS_ERR_ARG, "image smaller than the ROI");
    if (nframes == 0) return QCS_OK;
    hipStream_t s = (hipStream_t)stream;

    ExtractConst c;
    memset(&c, 0, sizeof(c));
    c.W = p.ImageWidth;
    c.H = p.ImageHigh;
    c.B = p.BatchedImageNum > 0 ? p.BatchedImageNum : 1;
    c.nframes = nframes;
    c.nbatch = (nframes + c.B - 1) / c.B;
    c.R = R;
    c.loc_type = p.LocType;
    c.multi = p.MultiEmitterFitEn;
    c.smooth = !(p.LocType == QCS_LOCTYPE_GS2D && p.MultiEmitterFitEn > 0);
    c.batch_cap = p.MaxROIPerBatch > 0 ? p.MaxROIPerBatch : QCS_MAX_POINT_NUM;
    c.cap = cap;
    c.start_frame = start_frame;
    c.nseg = (c.W + 63) / 64;
    filter_taps_host(R, p.LocType, p.MultiEmitterFitEn, c.taps);

    const size_t nrows = (size_t)nframes * c.H;
    if (nrows > (size_t)SCAN_TILE * SCAN_TILE) return fail(QCS_ERR_ARG, "too many rows in one extraction call");
    const int ntiles = (int)((nrows + SCAN_TILE - 1) / SCAN_TILE);
    int rc;
    if ((rc = ctx->ex_var.ensure(sizeof(unsigned long long) * 2 * c.nbatch)) ||
        (rc = ctx->ex_mark.ensure(sizeof(unsigned long long) * nrows * c.nseg * 2)) ||
        (rc = ctx->ex_cnt.ensure(sizeof(int) * (nrows * 2 + ntiles + 1))) ||
        (rc = ctx->ex_xy.ensure(sizeof(int3) * (size_t)(cap > 0 ? cap : 1))) ||
        (rc = ctx->ex_tmp.ensure(sizeof(int) * (2 + 3 * (size_t)c.nbatch))) ||
        (rc = ctx->pin_cnt.ensure(64)))
        return rc;
    unsigned long long *var_sums = ctx->ex_var.as<unsigned long long>();
    unsigned long long *cand = ctx->ex_mark.as<unsigned long long>();
    unsigned long long *mark = cand + nrows * c.nseg;
    int *row_count = ctx->ex_cnt.as<int>();
    int *row_off = row_count + nrows;
    int *tile_sums = row_off + nrows;
    int *info = ctx->ex_tmp.as<int>();
    int3 *xyb = ctx->ex_xy.as<int3>();

    hipEvent_t ev;
    if ((rc = timer_begin(ctx->timer, s, &ev))) return rc;
    fill_async(reinterpret_cast<int *>(var_sums), 4 * (size_t)c.nbatch, 0, s);
    fill_async(row_count, nrows, 0, s);
    hipLaunchKernelGGL(var_kernel, dim3(64, c.nbatch), dim3(256), 0, s, d_frames, c, var_sums);
    QCS_HIP(hipGetLastError());
    const int rows_per_batch = c.B * c.H;
    hipLaunchKernelGGL(detect_kernel, dim3(c.nseg, (rows_per_batch + TH - 1) / TH, c.nbatch), dim3(256), 0, s, d_frames, c,
                       var_sums, cand, mark, row_count);
    QCS_HIP(hipGetLastError());
    hipLaunchKernelGGL(scan_tiles_kernel, dim3(ntiles), dim3(SCAN_BLOCK), 0, s, row_count, (int)nrows, row_off, tile_sums);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(SCAN_BLOCK), 0, s, tile_sums, ntiles);
    hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, s, row_off, (int)nrows, tile_sums);
    hipLaunchKernelGGL(batch_kernel, dim3(1), dim3(64), 0, s, c, row_off, tile_sums, (int)nrows, ntiles, info);
    hipLaunchKernelGGL(emit_kernel, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, s, c, cand, row_off, row_count,
                       (int)nrows, info, xyb);
    QCS_HIP(hipGetLastError());
    if (cap > 0) {
        hipLaunchKernelGGL(cut_kernel, dim3((cap + 255) / 256), dim3(256), 0, s, d_frames, c, mark, info, xyb, d_rois);
        QCS_HIP(hipGetLastError());
        rc = wle_launch_devn(p, d_rois, info, cap, d_wle, s);
        if (rc) return rc;
    }
    if ((rc = timer_end(ctx->timer, TM_EXTRACT, s, ev, nframes))) return rc;
    int *h = ctx->pin_cnt.as<int>();
    QCS_HIP(hipMemcpyAsync(h, info, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipStreamSynchronize(s));
    if (n_out) *n_out = h[0];
    if (n_found) *n_found = h[1];
    if (h[1] > h[0]) return fail(QCS_ERR_CAPACITY, "%d candidates found, %d kept (capacity)", h[1], h[0]);
    return QCS_OK;
}

extern "C" int qcs_extract(qcs_ctx *ctx, const uint16_t *frames, int nframes, int start_frame, uint16_t *rois, float *wle,
                           int cap, int *n_out, int *n_found)
{
    if (n_out) *n_out = 0;
    if (n_found) *n_found = 0;
    if (!ctx || nframes < 0 || cap < 0 || (nframes > 0 && !frames) || (cap > 0 && (!rois || !wle)))
        return fail(QCS_ERR_ARG, "invalid argument");
    if (nframes == 0) return QCS_OK;
    const qcs_params &p = ctx->p;
    const size_t fb = (size_t)nframes * p.ImageWidth * p.ImageHigh * sizeof(uint16_t);
    const int R = p.ROISize;
    int rc;
    if ((rc = ctx->ex_frames.ensure(fb)) || (rc = ctx->h_rois_dev.ensure((size_t)(cap > 0 ? cap : 1) * QCS_ROI_WORDS(R) * 2)) ||
        (rc = ctx->h_wle_dev.ensure((size_t)(cap > 0 ? cap : 1) * QCS_WLE_FLOATS * 4)))
        return rc;
    QCS_HIP(hipMemcpy(ctx->ex_frames.ptr, frames, fb, hipMemcpyHostToDevice));
    int no = 0, nf = 0;
    rc = qcs_extract_dev(ctx, ctx->ex_frames.as<uint16_t>(), nframes, start_frame, ctx->h_rois_dev.as<uint16_t>(),
                         ctx->h_wle_dev.as<float>(), cap, &no, &nf, nullptr);
    if (rc && rc != QCS_ERR_CAPACITY) return rc;
    QCS_HIP(hipMemcpy(rois, ctx->h_rois_dev.ptr, (size_t)no * QCS_ROI_WORDS(R) * 2, hipMemcpyDeviceToHost));
    QCS_HIP(hipMemcpy(wle, ctx->h_wle_dev.ptr, (size_t)no * QCS_WLE_FLOATS * 4, hipMemcpyDeviceToHost));
    if (n_out) *n_out = no;
    if (n_found) *n_found = nf;
    return rc;
}
