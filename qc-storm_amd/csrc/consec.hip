// consec.hip -- consecutive-frame merging of localization rows on gfx950.
//
// Reference (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll):
//   ConsecutiveFit_WeightedAvg   DLL/ConsecutiveFit/ConsecutiveFit_top.cu:21-101
//   gpuFindConsecutiveMolecules  DLL/ConsecutiveFit/ConsecutiveFit.cu:18-77
//   gpuConsecutiveFit            DLL/ConsecutiveFit/ConsecutiveFit.cu:81-188
//   gpuRemoveConsecutiveFluo     DLL/ConsecutiveFit/ConsecutiveFit.cu:191-220
//
// A molecule that stays on over consecutive frames is localized once per frame; every row is
// linked to the closest row of the next frame within ConsecFit_DistanceTh_nm (integer-truncated
// positions, the first row in row order among equal distances), and each chain is replaced by its
// first row carrying the 10/CRLBx-weighted average position, z and PSF widths and the summed
// photons and background.  The rows of a call are held back until the next call, so chains that
// cross the boundary between calls are linked.
//
// The reference scans, for every row, all rows from its 32-row block start to the end of the next
// frame (O(rows x rows per frame)).  Here the rows are sorted once by (frame, cell) with cells
// wider than the distance threshold, and every row looks only at the 3x3 cells around its own in
// the next frame, keeping the (distance, row) minimum -- the same link as the reference's scan
// whenever the non-zero frames do not decrease along the rows (what frame-sorted compaction
// writes; rows zeroed by an earlier call have frame 0), which the call checks.  The merge walks
// and the removal walks are the reference's; the weighted sums are explicit FMAs where the
// reference build contracts `w += x * p`.
#include <limits.h>
#include <math.h>
#include <string.h>
#include <algorithm>
#include <new>
#include "qcs_internal.h"

namespace qcs {

namespace {

constexpr int CB = 32;        // the reference's scan starts at its 32-row block (ConsecutiveFit.h:31)
constexpr int ONT = QCS_ONTIME_BINS;
constexpr int FRAME_LIMIT = 1 << 30;

struct ConsecAcc {
    int flag;   // 1: frame out of [0, 2^30); 2: non-zero frames decrease along the rows
    int valid;
    int hist[ONT];
    int fmin, fmax, cymax, cxmax;   // key ranges of the call (frame clamped to [0, 2^30), cells)
};

// sort key (frame - fmin, cell row, cell column) packed with the call's field widths: every
// cell a link search visits (one beyond the largest present, at most 65535) and every frame
// up to the largest + 1 has its own key
struct KeyPack {
    int fmin, fshift, yshift;
};

// float -> int, truncating; NaN -> 0 and saturation (the conversion of the reference's device)
__device__ __forceinline__ int f2i(float x)
{
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT_MAX;
    if (x <= -2147483648.0f) return INT_MIN;
    return (int)x;
}

__device__ __forceinline__ int cell_of(int v, int cs)
{
    const long long q = v >= 0 ? (long long)v / cs : -((-(long long)v + cs - 1) / cs);
    return (int)(q < 0 ? 0 : (q > 65535 ? 65535 : q));
}

__device__ __forceinline__ unsigned long long cell_key(const KeyPack &k, int f, int cy, int cx)
{
    return ((unsigned long long)(unsigned)(f - k.fmin) << k.fshift) | ((unsigned long long)(unsigned)cy << k.yshift) |
           (unsigned long long)(unsigned)cx;
}

__global__ void consec_clear_kernel(ConsecAcc *acc)
{
    int *p = reinterpret_cast<int *>(acc);
    for (int k = threadIdx.x; k < (int)(sizeof(ConsecAcc) / sizeof(int)); k += blockDim.x) p[k] = 0;
    __syncthreads();
    if (threadIdx.x == 0) acc->fmin = INT_MAX;
}

// frames (clamped, as the keys use them) and cells: their ranges in acc.  Grid-stride over the
// rows (at most RANGE_BLOCKS work-groups), reduced per work-group, one atomic per range and group.
constexpr int RANGE_BLOCKS = 512;

__global__ void __launch_bounds__(256) consec_range_kernel(const float *__restrict__ a, int N, int cs, ConsecAcc *__restrict__ acc)
{
    __shared__ int red[4][4];
    int fmin = INT_MAX, fmax = 0, cymax = 0, cxmax = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const float *r = a + (size_t)i * QCS_REC_FLOATS;
        const int f = min(max(f2i(r[QCS_POS_FRME]), 0), FRAME_LIMIT - 1);
        fmin = min(fmin, f);
        fmax = max(fmax, f);
        cymax = max(cymax, cell_of(f2i(r[QCS_POS_YPOS]), cs));
        cxmax = max(cxmax, cell_of(f2i(r[QCS_POS_XPOS]), cs));
    }
    for (int off = 32; off > 0; off >>= 1) {
        fmin = min(fmin, __shfl_xor(fmin, off, 64));
        fmax = max(fmax, __shfl_xor(fmax, off, 64));
        cymax = max(cymax, __shfl_xor(cymax, off, 64));
        cxmax = max(cxmax, __shfl_xor(cxmax, off, 64));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[w][0] = fmin;
        red[w][1] = fmax;
        red[w][2] = cymax;
        red[w][3] = cxmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) {
            fmin = min(fmin, red[k][0]);
            fmax = max(fmax, red[k][1]);
            cymax = max(cymax, red[k][2]);
            cxmax = max(cxmax, red[k][3]);
        }
        atomicMin(&acc->fmin, fmin);
        atomicMax(&acc->fmax, fmax);
        atomicMax(&acc->cymax, cymax);
        atomicMax(&acc->cxmax, cxmax);
    }
}

// keys (frame, cell row, cell column), identity values, frames; clears the forward flags
__global__ void __launch_bounds__(256) consec_keys_kernel(const float *__restrict__ a, int N, int cs, KeyPack kp,
                                                          unsigned long long *__restrict__ keys, int *__restrict__ vals,
                                                          int *__restrict__ fr, int *__restrict__ fwd, ConsecAcc *__restrict__ acc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float *r = a + (size_t)i * QCS_REC_FLOATS;
    const int f = f2i(r[QCS_POS_FRME]);
    const int x = f2i(r[QCS_POS_XPOS]), y = f2i(r[QCS_POS_YPOS]);
    if (f < 0 || f >= FRAME_LIMIT) atomicOr(&acc->flag, 1);
    const int fc = min(max(f, 0), FRAME_LIMIT - 1);
    keys[i] = cell_key(kp, fc, cell_of(y, cs), cell_of(x, cs));
    vals[i] = i;
    fr[i] = fc;
    fwd[i] = 0;
}

__device__ __forceinline__ int lower_bound_u64(const unsigned long long *__restrict__ k, int n, unsigned long long v)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (k[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// gpuFindConsecutiveMolecules: the closest row of the next frame, ties to the lower row index
__global__ void __launch_bounds__(256) consec_link_kernel(const float *__restrict__ a, int N, int cs, KeyPack kp, float th2,
                                                          const unsigned long long *__restrict__ keys,
                                                          const int *__restrict__ vals, const int *__restrict__ fr,
                                                          const int *__restrict__ pmax, int *__restrict__ back,
                                                          int *__restrict__ fwd, ConsecAcc *__restrict__ acc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int f = fr[i];
    if (f != 0 && i > 0 && pmax[i - 1] > f) atomicOr(&acc->flag, 2);
    const float *r = a + (size_t)i * QCS_REC_FLOATS;
    const int x = f2i(r[QCS_POS_XPOS]), y = f2i(r[QCS_POS_YPOS]);
    const int cx = cell_of(x, cs), cy = cell_of(y, cs);
    const int j0 = i & ~(CB - 1);
    float best_d = 0.0f;
    int best = -1;
    if (f + 1 < FRAME_LIMIT) {
        for (int yy = max(cy - 1, 0); yy <= min(cy + 1, 65535); yy++)
            for (int xx = max(cx - 1, 0); xx <= min(cx + 1, 65535); xx++) {
                const unsigned long long k = cell_key(kp, f + 1, yy, xx);
                for (int q = lower_bound_u64(keys, N, k); q < N && keys[q] == k; q++) {
                    const int j = vals[q];
                    if (j < j0) continue;
                    const float *o = a + (size_t)j * QCS_REC_FLOATS;
                    const unsigned dx = (unsigned)x - (unsigned)f2i(o[QCS_POS_XPOS]);
                    const unsigned dy = (unsigned)y - (unsigned)f2i(o[QCS_POS_YPOS]);
                    const float d = (float)(int)(dx * dx + dy * dy);
                    if (d < th2 && (best < 0 || d < best_d || (d == best_d && j < best))) {
                        best = j;
                        best_d = d;
                    }
                }
            }
    }
    back[i] = best >= 0 ? best : -2;
    if (best >= 0) fwd[best] = 1;
}

// gpuConsecutiveFit: each chain head (no row links to it) takes the weighted average of its chain;
// other rows are copied.  On-time histogram over the heads of this call's new rows (i >= L).
__global__ void __launch_bounds__(256) consec_merge_kernel(const float *__restrict__ a, int N, int L, float sqrt_qe,
                                                           const int *__restrict__ back, const int *__restrict__ fwd,
                                                           float *__restrict__ out, ConsecAcc *__restrict__ acc)
{
    __shared__ int h[ONT];
    __shared__ int nvalid;
    if (threadIdx.x < ONT) h[threadIdx.x] = 0;
    if (threadIdx.x == 0) nvalid = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) {
        const float *src = a + (size_t)i * QCS_REC_FLOATS;
        float *dst = out + (size_t)i * QCS_REC_FLOATS;
        if (fwd[i] != 0) {
#pragma unroll
            for (int k = 0; k < QCS_REC_FLOATS; k++) dst[k] = src[k];
        } else {
            float wx = 0.0f, wy = 0.0f, wz = 0.0f, wsx = 0.0f, wsy = 0.0f, sp = 0.0f, st = 0.0f, sb = 0.0f, sw = 0.0f;
            int cur = i, cnt = 0;
            bool next = true;
            while (next && cnt < N) {   // chains follow strictly increasing frames: at most N rows
                const float *c = a + (size_t)cur * QCS_REC_FLOATS;
                const float prec = 10.0f / c[QCS_POS_CRBX];
                sp = sp + c[QCS_POS_PPHO];
                st = st + c[QCS_POS_TPHO];
                sb = sb + c[QCS_POS_BAKG];
                wx = __builtin_fmaf(c[QCS_POS_XPOS], prec, wx);
                wy = __builtin_fmaf(c[QCS_POS_YPOS], prec, wy);
                wz = __builtin_fmaf(c[QCS_POS_ZPOS], prec, wz);
                wsx = __builtin_fmaf(c[QCS_POS_SIGX], prec, wsx);
                wsy = __builtin_fmaf(c[QCS_POS_SIGY], prec, wsy);
                sw = sw + prec;
                cur = back[cur];
                next = cur > 0;
                cnt++;
            }
            dst[QCS_POS_PPHO] = sp;
            dst[QCS_POS_XPOS] = wx / sw;
            dst[QCS_POS_YPOS] = wy / sw;
            dst[QCS_POS_ZPOS] = wz / sw;
            dst[QCS_POS_SIGX] = wsx / sw;
            dst[QCS_POS_SIGY] = wsy / sw;
            dst[QCS_POS_TPHO] = st;
            dst[QCS_POS_BAKG] = sb;
            dst[QCS_POS_PSNR] = sqrt_qe * sp / sqrtf(sp + sb);
            dst[QCS_POS_CRBX] = src[QCS_POS_CRBX];
            dst[QCS_POS_CRBY] = src[QCS_POS_CRBY];
            dst[QCS_POS_FRME] = src[QCS_POS_FRME];
            if (i >= L) {
                atomicAdd(&h[min(max(cnt - 1, 0), ONT - 1)], 1);
                atomicAdd(&nvalid, 1);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < ONT && h[threadIdx.x]) atomicAdd(&acc->hist[threadIdx.x], h[threadIdx.x]);
    if (threadIdx.x == 0 && nvalid) atomicAdd(&acc->valid, nvalid);
}

// gpuRemoveConsecutiveFluo: every head zeroes the rest of its chain
__global__ void __launch_bounds__(256) consec_remove_kernel(int N, const int *__restrict__ back, const int *__restrict__ fwd,
                                                            float *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || fwd[i] != 0) return;
    int nx = back[i], steps = 0;
    while (nx > 0 && steps < N) {
        float *d = out + (size_t)nx * QCS_REC_FLOATS;
#pragma unroll
        for (int k = 0; k < QCS_REC_FLOATS; k++) d[k] = 0.0f;
        nx = back[nx];
        steps++;
    }
}

}  // namespace

}  // namespace qcs

using namespace qcs;

struct qcs_consec {
    qcs_ctx *ctx = nullptr;
    int n_last = 0;
    DevBuf last[2];   // held-back rows; the call writes the other buffer and swaps on success
    int cur = 0;
    DevBuf cat, mrg, keys, vals, fr, pmax, back, fwd, tmp, acc;
    HostBuf h_acc;
};

extern "C" {

int qcs_consec_create(qcs_ctx *ctx, qcs_consec **out)
{
    if (!ctx || !out) return fail(QCS_ERR_ARG, "invalid argument");
    *out = nullptr;
    qcs_consec *c = new (std::nothrow) qcs_consec();
    if (!c) return fail(QCS_ERR_NOMEM, "allocation failed");
    c->ctx = ctx;
    int rc;
    if ((rc = c->acc.ensure(sizeof(ConsecAcc))) || (rc = c->h_acc.ensure(sizeof(ConsecAcc)))) {
        qcs_consec_destroy(c);
        return rc;
    }
    *out = c;
    return QCS_OK;
}

int qcs_consec_held(const qcs_consec *c) { return c ? c->n_last : -QCS_ERR_ARG; }

int qcs_consec_reset(qcs_consec *c)
{
    if (!c) return fail(QCS_ERR_ARG, "invalid argument");
    c->n_last = 0;
    return QCS_OK;
}

int qcs_consec_destroy(qcs_consec *c)
{
    if (!c) return QCS_OK;
    DevBuf *b[] = {&c->last[0], &c->last[1], &c->cat, &c->mrg, &c->keys, &c->vals, &c->fr, &c->pmax, &c->back, &c->fwd,
                   &c->tmp, &c->acc};
    for (auto *x : b) x->release();
    c->h_acc.release();
    delete c;
    return QCS_OK;
}

int qcs_consec_fit_dev(qcs_consec *c, const float *d_rows, int n, int is_end, float *d_out, int out_cap, int *n_out,
                       float *ontime_ratio, void *stream)
{
    if (!c || n < 0 || out_cap < 0 || (n > 0 && !d_rows) || !n_out) return fail(QCS_ERR_ARG, "invalid argument");
    *n_out = 0;
    const qcs_params &p = c->ctx->p;
    hipStream_t s = (hipStream_t)stream;
    const int L = c->n_last;
    if ((long long)L + n > INT_MAX / QCS_REC_FLOATS) return fail(QCS_ERR_CAPACITY, "too many rows (%d held + %d)", L, n);
    const int N = L + n, out_n = is_end ? N : L;
    if (out_n > 0 && (!d_out || out_cap < out_n))
        return fail(QCS_ERR_CAPACITY, "out_cap (%d) < rows given back (%d)", out_cap, out_n);
    const size_t rb = (size_t)QCS_REC_FLOATS * sizeof(float);
    int rc;
    if (ontime_ratio)
        for (int k = 0; k < ONT; k++) ontime_ratio[k] = 0.0f;
    if (N == 0) {
        if (is_end) c->n_last = 0;
        return QCS_OK;
    }
    const size_t sort_b = sizeof(int) * radix_sort_tmp_ints(N), scan_b = sizeof(int) * max_scan_tmp_ints(N);
    const int nxt = c->cur ^ 1;
    if ((rc = c->cat.ensure((size_t)N * rb)) || (rc = c->mrg.ensure((size_t)N * rb)) ||
        (rc = c->keys.ensure((size_t)2 * N * 8)) || (rc = c->vals.ensure((size_t)2 * N * 4)) ||
        (rc = c->fr.ensure((size_t)N * 4)) || (rc = c->pmax.ensure((size_t)N * 4)) || (rc = c->back.ensure((size_t)N * 4)) ||
        (rc = c->fwd.ensure((size_t)N * 4)) || (rc = c->tmp.ensure(std::max(sort_b, scan_b) + 256)) ||
        (!is_end && n > 0 && (rc = c->last[nxt].ensure((size_t)n * rb))))
        return rc;
    float *cat = c->cat.as<float>(), *mrg = c->mrg.as<float>();
    if (L) QCS_HIP(hipMemcpyAsync(cat, c->last[c->cur].ptr, (size_t)L * rb, hipMemcpyDeviceToDevice, s));
    if (n) QCS_HIP(hipMemcpyAsync(cat + (size_t)L * QCS_REC_FLOATS, d_rows, (size_t)n * rb, hipMemcpyDeviceToDevice, s));

    const float th = p.ConsecFit_DistanceTh_nm / p.PixelSize;
    const float th2 = th * th;
    int cs = 1;
    if (th > 0.0f && th < 60000.0f) cs = (int)ceilf(th) + 1;
    unsigned long long *k_in = c->keys.as<unsigned long long>(), *k_out = k_in + N;
    int *v_in = c->vals.as<int>(), *v_out = v_in + N;
    ConsecAcc *acc = c->acc.as<ConsecAcc>();
    const dim3 g((N + 255) / 256), b(256);
    hipLaunchKernelGGL(consec_clear_kernel, dim3(1), dim3(64), 0, s, acc);
    hipLaunchKernelGGL(consec_range_kernel, dim3(std::min((N + 255) / 256, RANGE_BLOCKS)), b, 0, s, cat, N, cs, acc);
    QCS_HIP(hipGetLastError());
    // the key fields' widths from the call's ranges
    QCS_HIP(hipMemcpyAsync(c->h_acc.ptr, acc, sizeof(ConsecAcc), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipStreamSynchronize(s));
    const ConsecAcc *hr = c->h_acc.as<ConsecAcc>();
    auto nbits = [](long long v) { int b = 0; while (b < 62 && (v >> b) != 0) b++; return b; };
    KeyPack kp;
    kp.fmin = hr->fmin;
    const int xb = nbits(std::min(hr->cxmax + 1, 65535)), yb = nbits(std::min(hr->cymax + 1, 65535));
    const int fb = nbits((long long)hr->fmax + 1 - hr->fmin);
    kp.yshift = xb;
    kp.fshift = xb + yb;
    hipLaunchKernelGGL(consec_keys_kernel, g, b, 0, s, cat, N, cs, kp, k_in, v_in, c->fr.as<int>(), c->fwd.as<int>(), acc);
    QCS_HIP(hipGetLastError());
    unsigned long long *k_sorted;
    int *v_sorted;
    if ((rc = radix_sort_pairs_u64(k_in, v_in, k_out, v_out, N, fb + yb + xb, c->tmp.as<int>(), s, &k_sorted, &v_sorted)))
        return rc;
    if ((rc = scan_inclusive_max(c->fr.as<int>(), N, c->pmax.as<int>(), c->tmp.as<int>(), s))) return rc;
    hipLaunchKernelGGL(consec_link_kernel, g, b, 0, s, cat, N, cs, kp, th2, k_sorted, v_sorted, c->fr.as<int>(), c->pmax.as<int>(),
                       c->back.as<int>(), c->fwd.as<int>(), acc);
    QCS_HIP(hipGetLastError());
    hipLaunchKernelGGL(consec_merge_kernel, g, b, 0, s, cat, N, L, sqrtf(p.QE), c->back.as<int>(), c->fwd.as<int>(), mrg, acc);
    QCS_HIP(hipGetLastError());
    hipLaunchKernelGGL(consec_remove_kernel, g, b, 0, s, N, c->back.as<int>(), c->fwd.as<int>(), mrg);
    QCS_HIP(hipGetLastError());
    // CRLB of the rows held back from the previous call (LocPrecCalc_GaussianCRLB on FluoNum_LastGroup rows)
    if ((rc = crlb_launch(p, mrg, L, nullptr, s))) return rc;
    QCS_HIP(hipMemcpyAsync(c->h_acc.ptr, acc, sizeof(ConsecAcc), hipMemcpyDeviceToHost, s));
    if (out_n) QCS_HIP(hipMemcpyAsync(d_out, mrg, (size_t)out_n * rb, hipMemcpyDeviceToDevice, s));
    if (!is_end && n)
        QCS_HIP(hipMemcpyAsync(c->last[nxt].ptr, mrg + (size_t)L * QCS_REC_FLOATS, (size_t)n * rb, hipMemcpyDeviceToDevice, s));
    QCS_HIP(hipStreamSynchronize(s));
    const ConsecAcc *h = c->h_acc.as<ConsecAcc>();
    if (h->flag & 1) return fail(QCS_ERR_ARG, "consecutive merging: a frame number is outside [0, 2^30)");
    if (h->flag & 2)
        return fail(QCS_ERR_ARG, "consecutive merging: the rows are not sorted by frame (rows of frame 0 aside)");
    if (ontime_ratio) {
        const int v = h->valid <= 0 ? 1 : h->valid;
        for (int k = 0; k < ONT; k++) ontime_ratio[k] = (float)h->hist[k] * 1.0f / (float)v;
    }
    if (is_end) {
        c->n_last = 0;
    } else {
        c->n_last = n;
        c->cur = nxt;
    }
    *n_out = out_n;
    return QCS_OK;
}

}  // extern "C"
