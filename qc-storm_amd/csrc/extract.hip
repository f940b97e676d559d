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
//   1. walk_kernel<0>     background-removed values over each batch's variance region
//   2. var_reduce_kernel  exact integer sums of the variance samples per batch
//   3. walk_kernel<1>     one wave per 56-column strip walks down the rows: raw rows enter a
//                         9-row LDS ring, background removal, 5-tap H and V smoothing (uint16
//                         truncation after each pass, as the reference stores them), 3x3
//                         local-maximum test against the batch's thresholds.  Candidates and
//                         marks leave as one 56-bit ballot mask per row and strip; per-row
//                         candidate counts via integer atomics.
//   4. scan               exclusive prefix sum of the row counts -> raster-order offsets
//   5. batch_kernel       per-batch found / kept (capacity) / output offsets
//   6. emit_kernel        candidate (x, y) list in raster order
//   7. cut_kernel         detector classification (multi mode), raw R x R cut, header
//   8. WLE                PSF width, nearest neighbour (sorted sweep over the list), type
#include <string.h>
#include <type_traits>
#include "qcs_internal.h"

namespace qcs {

int wle_launch_extract(const qcs_params &p, const uint16_t *d_rois, const int3 *d_xyb, int B, int H, const int *d_n, int cap,
                       float *d_wle, hipStream_t s);
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
    int nseg;              // SEGW-column segments per row (one walking wave each)
    int walk_rows;         // image rows per walking wave
    float taps[5];
    const int *start_dev;  // when set: the call's first frame number is read here at run time (graph replays)
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// rows of batch b and its first frame
__device__ __forceinline__ int batch_rows(const ExtractConst &c, int b)
{
    const int f0 = b * c.B;
    const int nb = (c.nframes - f0) < c.B ? (c.nframes - f0) : c.B;
    return nb * c.H;
}

// variance region of batch b: S x S samples at (cx + i, cy + j), each the minimum of the
// background-removed image at (X, Y), (X - S, Y), (X, Y - S), (X - S, Y - S)
// (ImageVarianceCalc, SE/LDROIExtraction.cu:140-177; gpuImageVarianceCalcOnSelectRegion,
// SE/LDROIExtraction_ImageFiltering.cu:503-533)
__device__ __forceinline__ int var_side(const ExtractConst &c, int BH)
{
    int sw = c.W / 2 - 5, sh = BH / 2 - 5;
    sw = sw > 256 ? 256 : sw;
    sh = sh > 256 ? 256 : sh;
    sw = sw < 1 ? 1 : sw;
    sh = sh < 1 ? 1 : sh;
    return sw < sh ? sw : sh;
}

__device__ __forceinline__ void batch_thresholds(const ExtractConst &c, const unsigned long long *sums, int b, int BH,
                                                 float &th, float &th1)
{
    const int S = var_side(c, BH);
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

// ---- 1+2. row-walking filter wave ----------------------------------------------------------
// One wave (64 lanes, one column each) walks down a strip of the batched image.  Lane l owns
// column xc = x0 - 4 + l; the background sum needs raw columns xc-4..xc+4, so the raw rows in
// LDS are 72 wide (8 halo columns loaded by lanes 0..7).  Each stage narrows the valid lanes by
// its stencil reach: background sums on lanes 0..63, background-removed values on 1..62,
// H-smoothed values on 3..60, detection on 4..59 -- SEGW = 56 output columns per wave.
//
// A step handles 4 image rows r..r+3: the raw rows r+4..r+7 (loaded from HBM during the previous
// step) enter a 16-row LDS ring, then for each row r:
//   bk(r, x) = sum_{i<8} min(R(r-4, x-4+i), R(r-4+i, x+4), R(r+4, x+4-i), R(r+4-i, x-4))
//              with R the raw image, row/column indices clamped to the batch
//              (gpuRemoveBackgroundIntensity, SE/LDROIExtraction_ImageFiltering.cu:125-267);
//   tv(r, x) = uint16 trunc(max(raw - ((bk[xl] / 8 + bk[x] / 8) + bk[xr] / 8) / 3, 0)), xl/xr the
//              neighbours clamped to x's 32-column group.  The float expression equals the
//              integer max(raw - ceil((bk[xl] + bk[x] + bk[xr]) / 24), 0) exactly: the three
//              eighths sum exactly, s/3 rounds to within 2^-9 of S/24, and S/24 is either an
//              integer or at least 1/24 from one, so the truncation never changes;
//   hc(r, x) = uint16(min(fma-chain of tv(r, clamp(x-2+t)) * tap[t], 65535))      (H pass)
//   sv(y, x) = uint16(min(fma-chain of hc(clamp(y-2+t), x) * tap[t], 65535))      (V pass,
//              gpuImageConv_Seperable_H/V, SE/LDROIExtraction_ImageFiltering.cu:286-470);
//   detection of rows r-3..r once sv(r+1) exists (gpuROIFindingLD,
//   SE/LDROIExtraction_ROIExtraction.cu:25-164).
// Only the raw rows go through LDS (one barrier per step, after they are staged).  The other
// cross-lane values -- the neighbours' background sums, tv of lanes l-2..l+2 for the H pass and
// sv of lanes l-1, l+1 for detection -- move between lanes by DPP wave shifts, which keeps the
// step's chain of dependent LDS round trips short; the lanes whose results are used never reach
// past the wave, and a strip at the image border takes its clamped columns by readlane.  hc stays
// in a lane-local LDS ring, sv in registers.
// MODE 0 computes tv only, over the variance region, into tvbuf; MODE 1 is the full pass.
// Wave shifts by one lane (DPP wave_shr:1 / wave_shl:1).  Bound control writes 0 where the
// source lane is outside the wave (lane 0 for prev, lane 63 for next), so no old value is needed.
__device__ __forceinline__ int lane_prev(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int lane_next(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true); }
__device__ __forceinline__ float lane_prev(float v) { return __int_as_float(lane_prev(__float_as_int(v))); }
__device__ __forceinline__ float lane_next(float v) { return __int_as_float(lane_next(__float_as_int(v))); }

// two 16-bit lanes of a 32-bit register
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u16x2 as_u16x2(uint32_t w) { return __builtin_bit_cast(u16x2, w); }
__device__ __forceinline__ u16x2 swap16(u16x2 v) { return __builtin_shufflevector(v, v, 1, 0); }
__device__ __forceinline__ u16x2 pmin16(u16x2 a, u16x2 b) { return __builtin_elementwise_min(a, b); }
constexpr u16x2 kOnes16 = {1, 1};

constexpr int SEGW = 56;             // output columns per wave
constexpr int RAWW = 72;             // raw row width in LDS
constexpr int VAR_MAX = 512;         // region side bound (2 * 256)

constexpr int WALK_RS = 4;   // image rows per step
constexpr int WALK_NR = 16;  // raw ring rows (a step needs rows r-4 .. r+RS+3); power of two

// The walk of one strip.  BORDER: a lane of 1..62 has an H-pass column clamped to the image
// (the strips at the left and right image edges); only those carry the clamping code.
template <int MODE, bool SMOOTH, bool BORDER>
__device__ __forceinline__ void walk_body(const uint16_t *__restrict__ frames, const ExtractConst &c,
                                          const unsigned long long *__restrict__ var_sums,
                                          uint16_t *__restrict__ tvbuf,
                                          unsigned long long *__restrict__ cand_mask,
                                          unsigned long long *__restrict__ mark_mask,
                                          int *__restrict__ row_count, uint16_t (*raw)[RAWW], float (*hcs)[64])
{
    constexpr int RS = WALK_RS;
    constexpr int NR = WALK_NR;

    const int b = blockIdx.z;
    const int BH = batch_rows(c, b);
    const int seg = blockIdx.x;
    const int W = c.W;
    const uint16_t *img = frames + (size_t)b * c.B * c.H * W;
    const int l = threadIdx.x;

    int x0, ys, ye, rb, rlast;
    int S = 0, rx0 = 0, ry0 = 0;
    if (MODE == 0) {
        S = var_side(c, BH);
        rx0 = W / 2 - S;
        ry0 = BH / 2 - S;
        x0 = rx0 + blockIdx.x * SEGW;
        if (x0 >= rx0 + 2 * S) return;
        ys = ry0 + blockIdx.y * c.walk_rows;
        if (ys >= ry0 + 2 * S) return;
        ye = min(ys + c.walk_rows, ry0 + 2 * S);
        rb = ys;          // steps compute tv rows r .. r+3
        rlast = ye - 1;
    } else {
        x0 = seg * SEGW;
        ys = blockIdx.y * c.walk_rows;
        if (ys >= BH) return;
        ye = min(ys + c.walk_rows, BH);
        rb = ys - 4;      // step r: hc rows r..r+3, sv rows r-2..r+1, detection rows r-3..r
        rlast = ye + 2;
    }
    // ring slots are relative to rb; a step starts at rb + 4 * (4n + PH), so in interior steps the
    // slots are compile-time functions of the phase PH
    auto slot = [&](int v) { return (v - rb + 2 * NR) & (NR - 1); };
    auto ring8 = [&](int v) { return (v - rb + 16) & 7; };
    const int xc = x0 - 4 + l;
    const int xcl = clampi(xc, 0, W - 1);
    // halo columns: lanes 0..3 -> x0-8+l (slot l), lanes 4..7 -> x0+60+(l-4) (slot 68+l-4)
    const int hx = (l < 4) ? clampi(x0 - 8 + l, 0, W - 1) : clampi(x0 + 56 + l, 0, W - 1);
    const int hslot = (l < 4) ? l : 64 + l;
    auto ldp = [&](const uint16_t *src, uint16_t &main, uint16_t &halo) {
        main = src[xcl];
        halo = (l < 8) ? src[hx] : (uint16_t)0;
    };
    auto ld = [&](int v, uint16_t &main, uint16_t &halo) { ldp(img + (unsigned)(clampi(v, 0, BH - 1) * W), main, halo); };

    for (int v = rb - 4; v < rb + 4; v++) {
        uint16_t m, h;
        ld(v, m, h);
        raw[slot(v)][l + 4] = m;
        if (l < 8) raw[slot(v)][hslot] = h;
    }
    uint16_t pf[RS], pfh[RS];
#pragma unroll
    for (int k = 0; k < RS; k++) ld(rb + 4 + k, pf[k], pfh[k]);

    float th = 0.0f, th1 = 0.0f;
    float tap[5] = {0, 0, 0, 0, 0};
    if (MODE == 1) {
        batch_thresholds(c, var_sums, b, BH, th, th1);
#pragma unroll
        for (int t = 0; t < 5; t++) tap[t] = c.taps[t];
    }
    const int g0 = xc & ~31;
    const int li = max(min(((xc - 1 < g0) ? g0 : xc - 1) - (x0 - 4), 63), 0);
    const int ri = max(min(((xc + 1 > g0 + 31) ? g0 + 31 : xc + 1) - (x0 - 4), 63), 0);
    int hci[5];   // H-pass source lanes (columns clamped to the image)
#pragma unroll
    for (int t = 0; t < 5; t++) hci[t] = max(min(clampi(xc - 2 + t, 0, W - 1) - (x0 - 4), 63), 0);
    // lanes of image columns 0 and W-1 when the strip reaches a border (the H pass clamps to them)
    const int lane_lo = max(min(0 - (x0 - 4), 63), 0), lane_hi = max(min(W - 1 - (x0 - 4), 63), 0);
    const int half = c.R / 2;
    const bool out_lane = (l >= 4 && l < 4 + SEGW);
    const bool det_col = out_lane && xc < W && xc >= half && xc < W - half;
    const int ylo = max(ys - 1, 0), yhi = min(ye, BH - 1);   // sv rows this wave needs
    unsigned long long *cmask = cand_mask + (size_t)b * c.B * c.H * c.nseg + seg;
    unsigned long long *mmask = mark_mask + (size_t)b * c.B * c.H * c.nseg + seg;
    int *rcount = row_count + (size_t)b * c.B * c.H;
    const int nseg = c.nseg, H = c.H;
    int ryd = ((rb - 3) % H + H) % H;   // frame row of detection row r-3 for the current step
    // raw columns xc+4 (rows r-4 .. r+6) and xc-4 (rows r-3 .. r+7) of the current step; an
    // interior step after another step keeps the previous step's seven newest rows of each and
    // reads four from LDS (a walk's first step may be interior: then all eleven are read)
    int colR[RS + 7], colL[RS + 7];
    u16x2 CR[RS + 6], CL[RS + 6];
    bool have_cols = false;   // uniform
    // the lane's V-smoothed rows r-4 .. r+RS-3: [0, 2) from the previous step, [2, RS+2) this
    // step's V pass; detection takes lanes l-1 and l+1 of them by DPP wave shifts
    float svr[RS + 2];
    int lor[RS + 2];
#pragma unroll
    for (int i = 0; i < RS + 2; i++) { svr[i] = 0.0f; lor[i] = 0; }

    // One step of 4 rows.  EDGE: rows near the batch or walk ends (bounds checks, clamped rows,
    // dynamic ring slots); otherwise every row is in range and PH fixes the ring slots.
    auto step = [&](int r, auto edge_tag, auto ph_tag) {
        constexpr bool EDGE = decltype(edge_tag)::value;
        constexpr int PH = decltype(ph_tag)::value;
        auto rs = [&](int off) { if constexpr (EDGE) return slot(r + off); else return (4 * PH + off + 2 * NR) & (NR - 1); };
        auto h8 = [&](int off) { if constexpr (EDGE) return ring8(r + off); else return (4 * PH + off + 16) & 7; };
        // stage rows r+4 .. r+7 (loaded one step ago), then prefetch rows r+8 .. r+11
#pragma unroll
        for (int k = 0; k < RS; k++) {
            raw[rs(4 + k)][l + 4] = pf[k];
            if (l < 8) raw[rs(4 + k)][hslot] = pfh[k];
        }
        if constexpr (EDGE) {
#pragma unroll
            for (int k = 0; k < RS; k++) ld(r + 8 + k, pf[k], pfh[k]);
        } else {
            const uint16_t *p8 = img + (unsigned)((r + 8) * W);
#pragma unroll
            for (int k = 0; k < RS; k++) ldp(p8 + k * W, pf[k], pfh[k]);
        }
        __syncthreads();
        // background sums of rows r .. r+3
        // CR[m] / CL[m]: the 16-bit pairs (col[m], col[m+1]) of the two columns
        auto pack = [&](int m) {
            CR[m] = as_u16x2((uint32_t)colR[m] | ((uint32_t)colR[m + 1] << 16));
            CL[m] = as_u16x2((uint32_t)colL[m] | ((uint32_t)colL[m + 1] << 16));
        };
        if (EDGE || !have_cols) {
#pragma unroll
            for (int i = 0; i < RS + 7; i++) {
                colR[i] = raw[rs(i - 4)][l + 8];
                colL[i] = raw[rs(i - 3)][l];
            }
#pragma unroll
            for (int m = 0; m < RS + 6; m++) pack(m);
            have_cols = true;
        } else {
#pragma unroll
            for (int i = 0; i < 7; i++) {
                colR[i] = colR[i + RS];
                colL[i] = colL[i + RS];
            }
#pragma unroll
            for (int i = 7; i < RS + 7; i++) {
                colR[i] = raw[rs(i - 4)][l + 8];
                colL[i] = raw[rs(i - 3)][l];
            }
#pragma unroll
            for (int m = 0; m < 6; m++) {
                CR[m] = CR[m + RS];
                CL[m] = CL[m + RS];
            }
#pragma unroll
            for (int m = 6; m < RS + 6; m++) pack(m);
        }
        // Two terms per instruction: term i takes rA[l+i], colR[k+i], rB[l+8-i] and colL[k+7-i],
        // so terms 2j and 2j+1 read the 16-bit pairs (rA[l+2j], rA[l+2j+1]) and
        // (colR[k+2j], colR[k+2j+1]) in order and (rB[l+7-2j], rB[l+8-2j]) and
        // (colL[k+6-2j], colL[k+7-2j]) with their halves swapped.  Packed 16-bit minima, then a
        // dot product with (1, 1) adds both halves into the 32-bit sum.
        int bkv[RS];
#pragma unroll
        for (int k = 0; k < RS; k++) {
            uint4 A, B;   // rA[l .. l+7], rB[l+1 .. l+8]
            __builtin_memcpy(&A, &raw[rs(k - 4)][l], sizeof A);
            __builtin_memcpy(&B, &raw[rs(k + 4)][l + 1], sizeof B);
            const uint32_t aw[4] = {A.x, A.y, A.z, A.w}, bw[4] = {B.w, B.z, B.y, B.x};
            uint32_t bk = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const u16x2 m1 = pmin16(as_u16x2(aw[j]), CR[k + 2 * j]);
                const u16x2 m2 = pmin16(swap16(as_u16x2(bw[j])), swap16(CL[k + 6 - 2 * j]));
                bk = __builtin_amdgcn_udot2(pmin16(m1, m2), kOnes16, bk, false);
            }
            bkv[k] = (int)bk;
        }
        // the neighbours' sums by DPP wave shifts: li is l - 1 or l, ri is l or l + 1
        int tvk[RS];
#pragma unroll
        for (int k = 0; k < RS; k++) {
            const int bm = lane_prev(bkv[k]), bp = lane_next(bkv[k]);
            const int bl = li == l ? bkv[k] : bm, br = ri == l ? bkv[k] : bp;
            const int rawc = raw[rs(k)][l + 4];
            const int t = rawc - (bl + bkv[k] + br + 23) / 24;
            tvk[k] = t > 0 ? t : 0;
        }
        if constexpr (MODE == 0) {
#pragma unroll
            for (int k = 0; k < RS; k++) {
                const int v = r + k;
                if (out_lane && v < ye && xc < rx0 + 2 * S)
                    tvbuf[(size_t)b * VAR_MAX * VAR_MAX + (size_t)(v - ry0) * VAR_MAX + (xc - rx0)] = (uint16_t)tvk[k];
            }
            return;
        }
        // H pass into the lane's ring (rows inside the batch only).  The values of lanes l-2 .. l+2
        // come by DPP wave shifts; a strip at the image border takes its clamped columns from the
        // border lane (readlane).  Lanes 3 .. 60, the ones whose results are used, never read a
        // lane outside the wave.
#pragma unroll
        for (int k = 0; k < RS; k++) {
            if (EDGE && (r + k < 0 || r + k >= BH)) continue;   // uniform
            float hc;
            if constexpr (SMOOTH) {
                const float v0 = (float)tvk[k];
                const float m1 = lane_prev(v0), m2 = lane_prev(m1);   // lanes l-1, l-2
                const float p1 = lane_next(v0), p2 = lane_next(p1);   // lanes l+1, l+2
                float src[5] = {m2, m1, v0, p1, p2};
                if constexpr (BORDER) {
                    const float vl = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v0), lane_lo));
                    const float vr = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v0), lane_hi));
#pragma unroll
                    for (int t = 0; t < 5; t++) src[t] = hci[t] == lane_lo ? vl : (hci[t] == lane_hi ? vr : src[t]);
                }
                float acc = 0.0f;
#pragma unroll
                for (int t = 0; t < 5; t++) acc = __builtin_fmaf(src[t], tap[t], acc);
                acc = acc > 65535.0f ? 65535.0f : acc;
                hc = __builtin_truncf(acc);   // acc >= 0: the uint16 store's truncation
            } else {
                hc = (float)tvk[k];
            }
            hcs[h8(k)][l] = hc;
        }
        // V pass: sv rows r-2 .. r+1 (hc rows clamped to the batch), held as integral floats, and
        // whether each is at most th1
#pragma unroll
        for (int k = 0; k < RS; k++) {
            const int y = r - 2 + k;
            if (EDGE && (y < ylo || y > yhi)) continue;   // uniform
            float sv;
            if constexpr (SMOOTH) {
                float acc = 0.0f;
#pragma unroll
                for (int t = 0; t < 5; t++) {
                    float hv;
                    if constexpr (EDGE) hv = hcs[ring8(clampi(y - 2 + t, 0, BH - 1))][l];
                    else hv = hcs[h8(k - 4 + t)][l];
                    acc = __builtin_fmaf(hv, tap[t], acc);
                }
                acc = acc > 65535.0f ? 65535.0f : acc;
                sv = __builtin_truncf(acc);
            } else {
                sv = hcs[h8(k - 2)][l];
            }
            svr[2 + k] = sv;
            lor[2 + k] = (sv > th1) ? 0 : 1;
        }
        // detection rows r-3 .. r (sv rows r-4 .. r+1); the shifts run on the whole wave, ahead
        // of the per-lane tests.  The eight-neighbour count s8 >= 7 of the reference is "at most one
        // neighbour <= th1": with cv >= th > th1 the centre is not one of them, so it is the 3 x 3
        // count of the low flags (column sums, then lanes l-1, l, l+1) being at most one.
#pragma unroll
        for (int k = 0; k < RS; k++) {
            const int d = r - 3 + k;
            int ry = ryd + k;
            ry = ry >= H ? ry - H : ry;
            if (EDGE && (d < ys || d >= ye)) continue;   // uniform
            const float p01 = svr[k], cv = svr[k + 1], p21 = svr[k + 2];
            const float p00 = lane_prev(p01), p02 = lane_next(p01);
            const float p10 = lane_prev(cv), p12 = lane_next(cv);
            const float p20 = lane_prev(p21), p22 = lane_next(p21);
            const int lc = lor[k] + lor[k + 1] + lor[k + 2];
            const int low9 = lane_prev(lc) + lc + lane_next(lc);
            const bool sin4 = (cv > p01) && (cv >= p12) && (cv >= p21) && (cv > p10);
            const bool a = cv > p00, bb = cv > p02, cc = cv > p20, dd = cv > p22;
            const bool sout3 = (a && bb && (cc || dd)) || (cc && dd && (a || bb));
            const bool mark = det_col && ry >= half && ry < H - half && cv >= th && low9 <= 1 && sin4 && sout3;
            const bool cand = mark && a && bb && cc && dd;
            const unsigned long long cm = (__ballot(cand) >> 4) & ((1ULL << SEGW) - 1);
            const unsigned long long mm = (__ballot(mark) >> 4) & ((1ULL << SEGW) - 1);
            if (l == 0) {
                cmask[(size_t)d * nseg] = cm;
                mmask[(size_t)d * nseg] = mm;
                if (cm) atomicAdd(&rcount[d], __popcll(cm));
            }
        }
        svr[0] = svr[RS];
        svr[1] = svr[RS + 1];
        lor[0] = lor[RS];
        lor[1] = lor[RS + 1];
    };

    using E = std::true_type;
    using I = std::false_type;
    for (int r = rb; r <= rlast; r += RS) {
        bool edge;
        if (MODE == 0) edge = (r - 4 < 0) || (r + 11 > BH - 1) || (r + RS - 1 >= ye);
        else edge = (r - 4 < 0) || (r + 11 > BH - 1) || (r - 2 < ylo) || (r + 1 > yhi) || (r - 3 < ys) || (r >= ye);
        if (edge) {
            step(r, E{}, std::integral_constant<int, 0>{});
        } else {
            switch (((r - rb) >> 2) & 3) {
            case 0: step(r, I{}, std::integral_constant<int, 0>{}); break;
            case 1: step(r, I{}, std::integral_constant<int, 1>{}); break;
            case 2: step(r, I{}, std::integral_constant<int, 2>{}); break;
            default: step(r, I{}, std::integral_constant<int, 3>{}); break;
            }
        }
        ryd += RS;
        ryd = ryd >= H ? ryd - H : ryd;
    }
}

// The full pass is held to five waves per SIMD (at most 102 VGPRs, no spills): its steps are chains
// of dependent LDS round trips, and the fifth wave hides a little more of them (1% faster than
// four; six spill).
template <int MODE, bool SMOOTH>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MODE == 1 ? 5 : 1))) walk_kernel(const uint16_t *__restrict__ frames, ExtractConst c,
                                                  const unsigned long long *__restrict__ var_sums,
                                                  uint16_t *__restrict__ tvbuf,
                                                  unsigned long long *__restrict__ cand_mask,
                                                  unsigned long long *__restrict__ mark_mask,
                                                  int *__restrict__ row_count)
{
    __shared__ uint16_t raw[WALK_NR][RAWW];
    __shared__ float hcs[8][64];   // lane-local ring of H-smoothed rows
    const int xs = (int)blockIdx.x * SEGW;
    if (MODE == 1 && (xs - 4 + 1 < 2 || xs - 4 + 62 > c.W - 3))   // uniform
        walk_body<MODE, SMOOTH, true>(frames, c, var_sums, tvbuf, cand_mask, mark_mask, row_count, raw, hcs);
    else
        walk_body<MODE, SMOOTH, false>(frames, c, var_sums, tvbuf, cand_mask, mark_mask, row_count, raw, hcs);
}

// integer sums of the variance samples of each batch from the region's background-removed values
__global__ void __launch_bounds__(256) var_reduce_kernel(ExtractConst c, const uint16_t *__restrict__ tvbuf,
                                                         unsigned long long *__restrict__ sums)
{
    const int b = blockIdx.y;
    const int BH = batch_rows(c, b);
    const int S = var_side(c, BH);
    const uint16_t *t = tvbuf + (size_t)b * VAR_MAX * VAR_MAX;
    unsigned long long s1 = 0, s2 = 0;
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < S * S; g += gridDim.x * blockDim.x) {
        const int i = g % S, j = g / S;   // sample (X, Y) = (cx + i, cy + j), region origin (cx - S, cy - S)
        int d1 = t[(size_t)(S + j) * VAR_MAX + S + i];
        const int d2 = t[(size_t)(S + j) * VAR_MAX + i];
        int d3 = t[(size_t)j * VAR_MAX + S + i];
        const int d4 = t[(size_t)j * VAR_MAX + i];
        d1 = d1 < d2 ? d1 : d2;
        d3 = d3 < d4 ? d3 : d4;
        d1 = d1 < d3 ? d1 : d3;
        s1 += (unsigned long long)d1;
        s2 += (unsigned long long)d1 * (unsigned long long)d1;
    }
    for (int off = 32; off > 0; off >>= 1) {
        s1 += __shfl_down(s1, off, 64);
        s2 += __shfl_down(s2, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&sums[2 * b + 0], s1);
        atomicAdd(&sums[2 * b + 1], s2);
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
// One wave per stacked row: lane s takes segment s's candidate mask (coalesced loads), and a
// wave prefix sum of the masks' popcounts gives each segment its first slot, so the row's
// candidates come out in raster order.
constexpr int EMIT_WAVES = 4;   // rows per work-group
__global__ void __launch_bounds__(64 * EMIT_WAVES) emit_kernel(ExtractConst c, const unsigned long long *__restrict__ cand_mask,
                                                               const int *__restrict__ row_off, const int *__restrict__ row_count,
                                                               int nrows, const int *__restrict__ info, int3 *__restrict__ xyb)
{
    const int row = blockIdx.x * EMIT_WAVES + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= nrows || row_count[row] == 0) return;   // uniform over the wave
    const int b = row / (c.B * c.H);
    const int r0 = b * c.B * c.H;
    const int kept = info[2 + 3 * b + 1], out0 = info[2 + 3 * b + 2];
    int base = row_off[row] - row_off[r0];
    const int y = row - r0;   // batch-relative row
    for (int s0 = 0; s0 < c.nseg; s0 += 64) {
        const int sgi = s0 + lane;
        unsigned long long m = (sgi < c.nseg) ? cand_mask[(size_t)row * c.nseg + sgi] : 0ull;
        const int cnt = __popcll(m);
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if (lane >= o) incl += v;
        }
        int local = base + incl - cnt;
        while (m) {
            const int bit = __ffsll((long long)m) - 1;
            m &= m - 1;
            if (local < kept) xyb[out0 + local] = make_int3(sgi * SEGW + bit, y, b);
            local++;
        }
        base += __shfl(incl, 63, 64);
    }
}

// ---- 6. classification + cut -----------------------------------------------------------------
__device__ __forceinline__ int mark_count(const unsigned long long *mark, int nseg, size_t row, int xa, int xb)
{
    // number of marks in [xa, xb] of one stacked row (segments of SEGW columns)
    int cnt = 0;
    for (int sgi = xa / SEGW; sgi <= xb / SEGW; sgi++) {
        unsigned long long m = mark[row * nseg + sgi];
        const int lo = sgi * SEGW;
        if (xa > lo) m &= ~0ULL << (xa - lo);
        if (xb < lo + SEGW - 1) m &= ~0ULL >> (63 - (xb - lo));
        cnt += __popcll(m);
    }
    return cnt;
}

// one wave per group of CUT_U consecutive ROI records (grid-stride over the kept count): lanes
// write consecutive uint16 words of each record, so the stores are contiguous; the CUT_U records'
// list entries and pixels are loaded before any is stored, so a wave keeps CUT_U independent
// memory round trips in flight instead of one
constexpr int CUT_U = 8;

// R is the cut size; with sel set, record j is candidate sel[j] for j < *sel_n (a set of the
// mixed-size extraction), else record i is candidate i for i < info[0]
__global__ void __launch_bounds__(64) cut_kernel(const uint16_t *__restrict__ frames, ExtractConst c, int R,
                                                 const unsigned long long *__restrict__ mark_mask,
                                                 const int *__restrict__ info, const int3 *__restrict__ xyb,
                                                 const int *__restrict__ sel, const int *__restrict__ sel_n,
                                                 uint16_t *__restrict__ rois)
{
    // the CUT_U records are assembled in LDS and leave as one contiguous block (16-byte stores
    // when the block is 16-byte aligned)
    __shared__ __align__(16) uint16_t stage[CUT_U * 17 * 18];
    const int n = sel ? *sel_n : info[0];
    const int half = R / 2, WR = R * (R + 1), RR = R * R;
    const int l = threadIdx.x;
    for (int i0 = blockIdx.x * CUT_U; i0 < n; i0 += gridDim.x * CUT_U) {
        int3 e[CUT_U];
        int ft[CUT_U];
#pragma unroll
        for (int j = 0; j < CUT_U; j++) {
            e[j] = (i0 + j < n) ? xyb[sel ? sel[i0 + j] : i0 + j] : make_int3(half, half, 0);
            ft[j] = QCS_ROITYPE_SINGLE;
        }
        if (c.multi) {
#pragma unroll
            for (int j = 0; j < CUT_U; j++) {
                // lanes 0..R-1 count the marks of one window row each
                const int x = e[j].x, y = e[j].y;
                const size_t rbase = (size_t)e[j].z * c.B * c.H;
                int all = 0, inner = 0;
                if (l < R && i0 + j < n) {
                    const size_t row = rbase + y - half + l;
                    all = mark_count(mark_mask, c.nseg, row, x - half, x + half);
                    if (l >= half - 1 && l <= half + 1) inner = mark_count(mark_mask, c.nseg, row, x - 1, x + 1);
                }
                int d = all - inner;
                for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off, 64);
                ft[j] = (d > 0) ? QCS_ROITYPE_MULTI : QCS_ROITYPE_SINGLE;
            }
        }
        for (int k = l; k < WR; k += 64) {
            uint16_t v[CUT_U];
#pragma unroll
            for (int j = 0; j < CUT_U; j++) {
                const int x = e[j].x, y = e[j].y, b = e[j].z;
                if (k < RR) {
                    const int r = k / R, cc = k - r * R;
                    const uint16_t *img = frames + (size_t)b * c.B * c.H * c.W;
                    v[j] = (i0 + j < n) ? img[(size_t)(y - half + r) * c.W + (x - half + cc)] : (uint16_t)0;
                } else {
                    const int h = k - RR;
                    const int fr = (c.start_dev ? *c.start_dev : c.start_frame) + b * c.B + y / c.H;
                    v[j] = h == 0 ? (uint16_t)x
                         : h == 1 ? (uint16_t)(y % c.H)
                         : h == 2 ? (uint16_t)(fr % 65536)
                         : h == 3 ? (uint16_t)(fr / 65536)
                         : h == 4 ? (uint16_t)ft[j] : (uint16_t)0;
                }
            }
#pragma unroll
            for (int j = 0; j < CUT_U; j++) stage[j * WR + k] = v[j];
        }
        __syncthreads();
        const int nv = min(CUT_U, n - i0);
        uint16_t *dst = rois + (size_t)i0 * WR;
        const uintptr_t al = reinterpret_cast<uintptr_t>(dst);
        if (nv == CUT_U && (al & 15) == 0) {   // uniform; CUT_U records are 16 R (R + 1) bytes
            for (int t = l; t < CUT_U * WR / 8; t += 64)
                reinterpret_cast<uint4 *>(dst)[t] = reinterpret_cast<const uint4 *>(stage)[t];
        } else if ((al & 3) == 0) {   // records are a whole number of 4-byte words (R (R + 1) even)
            for (int t = l; t < nv * WR / 2; t += 64)
                reinterpret_cast<uint32_t *>(dst)[t] = reinterpret_cast<const uint32_t *>(stage)[t];
        } else {
            for (int t = l; t < nv * WR; t += 64) dst[t] = stage[t];
        }
        __syncthreads();   // the next group reuses stage
    }
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

namespace qcs {

// Everything of one extraction call up to the candidate list (no host synchronisation):
// variance, walking filter + detection, row scan, per-batch bookkeeping, raster-order (x, y,
// batch) list.  info[0] receives the kept count (at most cap), info[1] the found count.
struct ExtractRun {
    ExtractConst c;
    int *info;
    int3 *xyb;
    const unsigned long long *mark;
};

static int extract_candidates(qcs_ctx *ctx, const qcs_params &p, const uint16_t *d_frames, int nframes, int start_frame,
                              const int *d_start, int cap, hipStream_t s, ExtractRun &run)
{
    const int R = p.ROISize;
    ExtractConst &c = run.c;
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
    c.start_dev = d_start;
    c.nseg = (c.W + SEGW - 1) / SEGW;
    filter_taps_host(R, p.LocType, p.MultiEmitterFitEn, c.taps);
    // rows per walking wave: long walks amortise the 14-row pipeline fill; shorter ones keep
    // small calls spread over the chip
    const int rows_per_batch = c.B * c.H;
    c.walk_rows = 128;
    while (c.walk_rows > 16 &&
           (long long)c.nseg * ((rows_per_batch + c.walk_rows - 1) / c.walk_rows) * c.nbatch < 8192)
        c.walk_rows /= 2;

    const size_t nrows = (size_t)nframes * c.H;
    if (nrows > (size_t)SCAN_TILE * SCAN_TILE) return fail(QCS_ERR_ARG, "too many rows in one extraction call");
    if ((size_t)rows_per_batch * c.W >= ((size_t)1 << 31)) return fail(QCS_ERR_ARG, "batched image larger than 2^31 pixels");
    const int ntiles = (int)((nrows + SCAN_TILE - 1) / SCAN_TILE);
    int rc;
    if ((rc = ctx->ex_var.ensure(sizeof(unsigned long long) * 2 * c.nbatch)) ||
        (rc = ctx->ex_det.ensure(sizeof(uint16_t) * VAR_MAX * VAR_MAX * (size_t)c.nbatch)) ||
        (rc = ctx->ex_mark.ensure(sizeof(unsigned long long) * nrows * c.nseg * 2)) ||
        (rc = ctx->ex_cnt.ensure(sizeof(int) * (nrows * 2 + ntiles + 1))) ||
        (rc = ctx->ex_xy.ensure(sizeof(int3) * (size_t)(cap > 0 ? cap : 1))) ||
        (rc = ctx->ex_tmp.ensure(sizeof(int) * (2 + 3 * (size_t)c.nbatch))))
        return rc;
    unsigned long long *var_sums = ctx->ex_var.as<unsigned long long>();
    uint16_t *tvbuf = ctx->ex_det.as<uint16_t>();
    unsigned long long *cand = ctx->ex_mark.as<unsigned long long>();
    unsigned long long *mark = cand + nrows * c.nseg;
    int *row_count = ctx->ex_cnt.as<int>();
    int *row_off = row_count + nrows;
    int *tile_sums = row_off + nrows;
    int *info = ctx->ex_tmp.as<int>();
    int3 *xyb = ctx->ex_xy.as<int3>();

    // variance region extent (largest over the batches: the first batch is a full one)
    int Smax;
    {
        int sw = c.W / 2 - 5, sh = rows_per_batch / 2 - 5;
        sw = sw > 256 ? 256 : (sw < 1 ? 1 : sw);
        sh = sh > 256 ? 256 : (sh < 1 ? 1 : sh);
        Smax = sw < sh ? sw : sh;
    }
    fill_async(reinterpret_cast<int *>(var_sums), 4 * (size_t)c.nbatch, 0, s);
    fill_async(row_count, nrows, 0, s);
    hipLaunchKernelGGL((walk_kernel<0, false>), dim3((2 * Smax + SEGW - 1) / SEGW, (2 * Smax + c.walk_rows - 1) / c.walk_rows, c.nbatch),
                       dim3(64), 0, s, d_frames, c, var_sums, tvbuf, cand, mark, row_count);
    QCS_HIP(hipGetLastError());
    hipLaunchKernelGGL(var_reduce_kernel, dim3(16, c.nbatch), dim3(256), 0, s, c, tvbuf, var_sums);
    QCS_HIP(hipGetLastError());
    const dim3 wgrid(c.nseg, (rows_per_batch + c.walk_rows - 1) / c.walk_rows, c.nbatch);
    if (c.smooth)
        hipLaunchKernelGGL((walk_kernel<1, true>), wgrid, dim3(64), 0, s, d_frames, c, var_sums, tvbuf, cand, mark, row_count);
    else
        hipLaunchKernelGGL((walk_kernel<1, false>), wgrid, dim3(64), 0, s, d_frames, c, var_sums, tvbuf, cand, mark, row_count);
    QCS_HIP(hipGetLastError());
    hipLaunchKernelGGL(scan_tiles_kernel, dim3(ntiles), dim3(SCAN_BLOCK), 0, s, row_count, (int)nrows, row_off, tile_sums);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(SCAN_BLOCK), 0, s, tile_sums, ntiles);
    hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, s, row_off, (int)nrows, tile_sums);
    hipLaunchKernelGGL(batch_kernel, dim3(1), dim3(64), 0, s, c, row_off, tile_sums, (int)nrows, ntiles, info);
    hipLaunchKernelGGL(emit_kernel, dim3((unsigned)((nrows + EMIT_WAVES - 1) / EMIT_WAVES)), dim3(64 * EMIT_WAVES), 0, s, c, cand,
                       row_off, row_count, (int)nrows, info, xyb);
    QCS_HIP(hipGetLastError());
    run.info = info;
    run.xyb = xyb;
    run.mark = mark;
    return QCS_OK;
}

static int check_extract_args(const qcs_params &p, int R)
{
    if (p.ImageWidth <= 0 || p.ImageHigh <= 0) return fail(QCS_ERR_ARG, "ImageWidth/ImageHigh not set");
    if (p.ImageWidth < R || p.ImageHigh < R) return fail(QCS_ERR_ARG, "image smaller than the ROI");
    return QCS_OK;
}

// ---- mixed ROI sizes (no reference counterpart) ----------------------------------------------
// Each candidate of the base-size-7 extraction gets an ROI size in {7, 9, 11}:
//   s    = min(SigmaL, SigmaR, SigmaU, SigmaD) of its 7 x 7 WLE record;
//   Rpsf = 7 if s <= 1.4, 9 if s <= 1.9, else 11 (about +-2.5 s);
//   c(d) = detector marks in the (2d + 1)^2 window, rows of the candidate's own frame only;
//   dmin = smallest d in 1..5 with c(d) > 1 (6 if none);
//   Rnb  = 2 dmin + 3 for dmin 3 or 4 (a neighbour on the edge of the 7 x 7 window is taken in
//          whole, with a pixel of margin), else 7;
//   R    = max(Rpsf, Rnb), reduced by 2 while R > 7 and the window holds more than 2 other marks
//          (a fit models at most 3 emitters), then while R > 7 and the window leaves the frame.
// flags[s * ncap + i] = 1 when candidate i gets size 7 + 2 s; entries i >= *n are cleared.
__global__ void __launch_bounds__(256) mixed_size_kernel(ExtractConst c, const unsigned long long *__restrict__ mark_mask,
                                                         const int *__restrict__ info, const int3 *__restrict__ xyb,
                                                         const float *__restrict__ wle7, int ncap, int *__restrict__ flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncap) return;
    int R = 0;
    if (i < info[0]) {
        const int3 e = xyb[i];
        const int x = e.x, y = e.y;                 // y: row of the stacked batch
        const int fy0 = (y / c.H) * c.H, ry = y - fy0;
        const size_t rbase = (size_t)e.z * c.B * c.H;
        const float *w = wle7 + (size_t)i * QCS_WLE_FLOATS;
        float sg = w[QCS_WLE_SIGMAL];
        sg = w[QCS_WLE_SIGMAR] < sg ? w[QCS_WLE_SIGMAR] : sg;
        sg = w[QCS_WLE_SIGMAU] < sg ? w[QCS_WLE_SIGMAU] : sg;
        sg = w[QCS_WLE_SIGMAD] < sg ? w[QCS_WLE_SIGMAD] : sg;
        const int rpsf = (sg <= 1.4f) ? 7 : ((sg <= 1.9f) ? 9 : 11);
        int cnt[6];
        cnt[0] = 1;
        for (int d = 1; d <= 5; d++) {
            const int xa = x - d < 0 ? 0 : x - d, xb = x + d > c.W - 1 ? c.W - 1 : x + d;
            int ya = y - d, yb = y + d;
            ya = ya < fy0 ? fy0 : ya;
            yb = yb > fy0 + c.H - 1 ? fy0 + c.H - 1 : yb;
            int n = 0;
            for (int yy = ya; yy <= yb; yy++) n += mark_count(mark_mask, c.nseg, rbase + yy, xa, xb);
            cnt[d] = n;
        }
        int dmin = 6;
        for (int d = 5; d >= 1; d--)
            if (cnt[d] > 1) dmin = d;
        const int rnb = (dmin == 3 || dmin == 4) ? 2 * dmin + 3 : 7;
        R = rpsf > rnb ? rpsf : rnb;
        while (R > 7 && cnt[R / 2] - 1 > 2) R -= 2;
        while (R > 7 && !(x - R / 2 >= 0 && x + R / 2 < c.W && ry - R / 2 >= 0 && ry + R / 2 < c.H)) R -= 2;
    }
    flags[i] = R == 7;
    flags[(size_t)ncap + i] = R == 9;
    flags[2 * (size_t)ncap + i] = R == 11;
}

// set lists: sel[off[i]] = i for flagged i with off[i] < cap; *n_out = min(total, cap)
__global__ void __launch_bounds__(256) mixed_scatter_kernel(const int *__restrict__ flags, const int *__restrict__ offs, int ncap,
                                                            int cap, int *__restrict__ sel)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncap || !flags[i]) return;
    const int o = offs[i];
    if (o < cap) sel[o] = i;
}

__global__ void mixed_count_kernel(const int *__restrict__ tile_sums, int ntiles, int cap, int *__restrict__ n_out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int t = tile_sums[ntiles];
    *n_out = t < cap ? t : cap;
}

// Mixed-size extraction (launches only): candidates of the base-size-7 extraction (at most cap),
// their 7 x 7 cuts and WLE records in ctx scratch, the size rule, and per set s (sizes 7, 9, 11)
// the candidate list cand_s (ctx scratch, returned), the count sets[s].d_n and the records
// sets[s].d_rois (at most sets[s].n).
int extract_mixed_launch(qcs_ctx *ctx, const uint16_t *d_frames, int nframes, int start_frame, const int *d_start, int cap,
                         const qcs_roi_set *sets, int *cand_lists[3], int **cand_info, hipStream_t s)
{
    qcs_params p = ctx->p;
    p.ROISize = 7;
    int rc;
    if ((rc = check_extract_args(p, 11))) return rc;
    for (int k = 0; k < 3; k++)
        if (sets[k].roi_size != 7 + 2 * k || sets[k].n < 0 || (sets[k].n > 0 && !sets[k].d_rois) || !sets[k].d_n)
            return fail(QCS_ERR_ARG, "set %d: sizes 7, 9, 11 in order, a record buffer and a device count are required", k);
    ExtractRun run;
    if ((rc = extract_candidates(ctx, p, d_frames, nframes, start_frame, d_start, cap, s, run))) return rc;
    const int ncap = cap > 0 ? cap : 1;
    const int ntiles = scan_tile_count((size_t)ncap);
    // scratch: 7 x 7 cuts, WLE records, flags (3 ncap), offsets (3 ncap), tile sums (3 (ntiles + 1)), lists (3 ncap)
    const size_t r7w = (size_t)QCS_ROI_WORDS(7) * ncap;
    if ((rc = ctx->mx_rois.ensure(sizeof(uint16_t) * r7w + sizeof(float) * QCS_WLE_FLOATS * (size_t)ncap)) ||
        (rc = ctx->mx_list.ensure(sizeof(int) * (9 * (size_t)ncap + 3 * ((size_t)ntiles + 1)))))
        return rc;
    uint16_t *r7 = ctx->mx_rois.as<uint16_t>();
    float *w7 = reinterpret_cast<float *>(r7 + r7w);
    int *flags = ctx->mx_list.as<int>();
    int *offs = flags + 3 * (size_t)ncap;
    int *lists = offs + 3 * (size_t)ncap;
    int *tsum = lists + 3 * (size_t)ncap;
    if (cap > 0) {
        const int cut_blocks = cap < 16384 ? cap : 16384;
        hipLaunchKernelGGL(cut_kernel, dim3(cut_blocks), dim3(64), 0, s, d_frames, run.c, 7, run.mark, run.info, run.xyb, nullptr, nullptr,
                           r7);
        QCS_HIP(hipGetLastError());
        if ((rc = wle_launch_extract(p, r7, run.xyb, run.c.B, run.c.H, run.info, cap, w7, s))) return rc;
    }
    hipLaunchKernelGGL(mixed_size_kernel, dim3((ncap + 255) / 256), dim3(256), 0, s, run.c, run.mark, run.info, run.xyb, w7, ncap, flags);
    QCS_HIP(hipGetLastError());
    for (int k = 0; k < 3; k++) {
        int *f = flags + (size_t)k * ncap, *o = offs + (size_t)k * ncap, *l = lists + (size_t)k * ncap;
        int *ts = tsum + (size_t)k * (ntiles + 1);
        if ((rc = scan_exclusive(f, ncap, o, ts, s))) return rc;
        hipLaunchKernelGGL(mixed_scatter_kernel, dim3((ncap + 255) / 256), dim3(256), 0, s, f, o, ncap, sets[k].n, l);
        hipLaunchKernelGGL(mixed_count_kernel, dim3(1), dim3(1), 0, s, ts, ntiles, sets[k].n, sets[k].d_n);
        QCS_HIP(hipGetLastError());
        if (sets[k].n > 0) {
            const int blocks = sets[k].n < 16384 ? sets[k].n : 16384;
            hipLaunchKernelGGL(cut_kernel, dim3(blocks), dim3(64), 0, s, d_frames, run.c, 7 + 2 * k, run.mark, run.info, run.xyb, l,
                               sets[k].d_n, sets[k].d_rois);
            QCS_HIP(hipGetLastError());
        }
        cand_lists[k] = l;
    }
    if (cand_info) *cand_info = run.info;
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
    int rc;
    if ((rc = check_extract_args(p, p.ROISize))) return rc;
    if (nframes == 0) return QCS_OK;
    hipStream_t s = (hipStream_t)stream;
    if ((rc = ctx->pin_cnt.ensure(64))) return rc;
    hipEvent_t ev;
    if ((rc = timer_begin(ctx->timer, s, &ev))) return rc;
    ExtractRun run;
    if ((rc = extract_candidates(ctx, p, d_frames, nframes, start_frame, nullptr, cap, s, run))) return rc;
    if (cap > 0) {
        const int cut_blocks = cap < 16384 ? cap : 16384;
        hipLaunchKernelGGL(cut_kernel, dim3(cut_blocks), dim3(64), 0, s, d_frames, run.c, run.c.R, run.mark, run.info, run.xyb, nullptr,
                           nullptr, d_rois);
        QCS_HIP(hipGetLastError());
        rc = wle_launch_extract(p, d_rois, run.xyb, run.c.B, run.c.H, run.info, cap, d_wle, s);
        if (rc) return rc;
    }
    if ((rc = timer_end(ctx->timer, TM_EXTRACT, s, ev, nframes))) return rc;
    int *h = ctx->pin_cnt.as<int>();
    QCS_HIP(hipMemcpyAsync(h, run.info, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipStreamSynchronize(s));
    if (n_out) *n_out = h[0];
    if (n_found) *n_found = h[1];
    if (h[1] > h[0]) return fail(QCS_ERR_CAPACITY, "%d candidates found, %d kept (capacity)", h[1], h[0]);
    return QCS_OK;
}

extern "C" int qcs_extract_mixed_dev(qcs_ctx *ctx, const uint16_t *d_frames, int nframes, int start_frame, int cap,
                                     const qcs_roi_set *sets, int *n_counts, void *stream)
{
    if (!ctx || !sets || nframes < 0 || cap < 0 || (nframes > 0 && !d_frames)) return fail(QCS_ERR_ARG, "invalid argument");
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if (nframes == 0) {
        for (int k = 0; k < 3; k++) {
            if (!sets[k].d_n) return fail(QCS_ERR_ARG, "set %d: a device count is required", k);
            fill_async(sets[k].d_n, 1, 0, s);
        }
        QCS_HIP(hipGetLastError());
    } else {
        hipEvent_t ev;
        if ((rc = timer_begin(ctx->timer, s, &ev))) return rc;
        int *lists[3], *info = nullptr;
        if ((rc = extract_mixed_launch(ctx, d_frames, nframes, start_frame, nullptr, cap, sets, lists, &info, s))) return rc;
        if ((rc = timer_end(ctx->timer, TM_EXTRACT, s, ev, nframes))) return rc;
    }
    if (!n_counts) return QCS_OK;
    int h[3];
    for (int k = 0; k < 3; k++) QCS_HIP(hipMemcpyAsync(&h[k], sets[k].d_n, sizeof(int), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < 3; k++) n_counts[k] = h[k];
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
