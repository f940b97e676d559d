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
// in row order (scan), then stably radix-sorted by frame (sort.hip), so rows of one frame keep their
// original order.  Rendering splats each localization's 7 x 7 Gaussian with float atomics, as the
// reference does, but into LDS: localizations are binned by 64 x 64 tile of the SR image (a
// footprint that straddles a tile edge is listed in each tile it touches), one work-group
// accumulates a tile with LDS atomics and then adds it to the image with one coalesced
// read-modify-write.  Scattered atomics into a 1.7 GB image would miss the caches on nearly
// every update.  The PSF uses the deterministic exponential.
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
    int tiles_x, tiles_y;
};
constexpr int RTILE = 64;   // SR tile edge, pixels

// The drawing parameters of one localization row (FluoRenderWithPrec_2D/_3D,
// ImageRender_2d.cu:55-155, ImageRender_3d.cu:60-147): false if it is not drawn.
struct Splat {
    float pp, rs, fx, fy, z;
    int ix, iy;
    bool one;   // a single pixel (peak photons, 2D)
};
__device__ __forceinline__ bool render_prep(const float *__restrict__ r, const RenderConst &c, Splat &o)
{
    float pp = r[QCS_POS_PPHO];
    const float rx = r[QCS_POS_XPOS] * c.zoom, ry = r[QCS_POS_YPOS] * c.zoom;
    float snr = r[QCS_POS_PSNR];
    float lp = c.is3d ? (r[QCS_POS_CRBX] + r[QCS_POS_CRBY]) / 2.0f : r[QCS_POS_CRBX];
    o.z = r[QCS_POS_ZPOS];
    if (lp != lp) { snr = 0.0f; pp = 0.0f; }
    if (lp <= 0.0f) pp = 0.0f;
    bool inside = snr >= c.snr_th && pp > 1.0f && rx > 7.0f && rx < (float)(c.srw - 1 - 7) && ry > 7.0f &&
                  ry < (float)(c.srh - 1 - 7);
    if (c.is3d) inside = inside && fabsf(o.z) < c.zth;
    if (!inside) return false;
    if (c.mode == QCS_RENDER_1PHOTON_CRLB || c.mode == QCS_RENDER_1PHOTON_FIXED) pp = 1.0f;
    if (c.mode == QCS_RENDER_1PHOTON_FIXED) lp = c.fixed_prec;
    if (!c.is3d && lp > 30.0f) { lp = 30.0f; pp = 0.0f; }
    lp = fmaxf(lp, (c.ps / c.zoom) / 2.0f);
    float rs = lp / (c.ps / c.zoom);
    o.rs = -0.5f / (rs * rs);
    o.pp = pp;
    o.ix = (int)rx;
    o.iy = (int)ry;
    o.fx = (rx - (float)o.ix) + 3.0f;
    o.fy = (ry - (float)o.iy) + 3.0f;
    o.one = !c.is3d && c.mode == QCS_RENDER_PEAKPHOTON_1PIXEL;
    return true;
}

// tiles touched by the footprint of a drawn localization
__device__ __forceinline__ void splat_tiles(const Splat &q, int &tx0, int &tx1, int &ty0, int &ty1)
{
    const int h = q.one ? 0 : 3;
    tx0 = (q.ix - h) / RTILE;
    tx1 = (q.ix + h) / RTILE;
    ty0 = (q.iy - h) / RTILE;
    ty1 = (q.iy + h) / RTILE;
}

// A localization drawn into one tile, as the tile kernel reads it: two 16-byte words,
// {pp, rs, fx, fy} and {z, (iy - y0) << 16 | (ix - x0) & 0xffff (tile-relative), 0, 0}.
constexpr int RENT = 2;   // float4 words per entry

// pass 1 (cursor == null): localizations per tile; pass 2: the per-tile lists of prepared entries
__global__ void __launch_bounds__(256) render_bin_kernel(const float *__restrict__ loc, int n, RenderConst c,
                                                         int *__restrict__ tcnt, const int *__restrict__ toff,
                                                         int *__restrict__ cursor, float4 *__restrict__ ent)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Splat q;
    if (!render_prep(loc + (size_t)i * QCS_REC_FLOATS, c, q)) return;
    int tx0, tx1, ty0, ty1;
    splat_tiles(q, tx0, tx1, ty0, ty1);
    for (int ty = ty0; ty <= ty1; ty++)
        for (int tx = tx0; tx <= tx1; tx++) {
            const int t = ty * c.tiles_x + tx;
            if (!cursor) {
                atomicAdd(&tcnt[t], 1);
                continue;
            }
            const size_t e = (size_t)(toff[t] + atomicAdd(&cursor[t], 1)) * RENT;
            const int rx = q.ix - tx * RTILE, ry = q.iy - ty * RTILE;   // within [-3, RTILE + 2]
            ent[e] = make_float4(q.pp, q.rs, q.fx, q.fy);
            ent[e + 1] = make_float4(q.z, __int_as_float((int)(((unsigned)ry << 16) | ((unsigned)rx & 0xffffu))), 0.0f, 0.0f);
        }
}

// One work-group of RBLK threads per tile: LDS accumulation, then img += tile (and zimg += z tile
// with ZON).  The tile's entries are contiguous; each thread stages one in LDS per round, then
// each wave draws one staged localization at a time, one footprint pixel per lane.  An interior tile of an aligned image (rows 16-byte aligned,
// width a multiple of 4) reads its image pixels, four per thread, before drawing, so the HBM
// latency of the read-modify-write overlaps the LDS work.  FRESH: the planes' previous contents are
// replaced (as if cleared first): every tile is written, empty ones with zeros, and nothing is
// read, so a fresh image costs one write of the planes instead of a clear plus a read-modify-write.
constexpr int RBLK = 256;
constexpr int RVEC = RTILE * RTILE / (4 * RBLK);   // float4 image words per thread
constexpr int RSTR = RTILE;   // LDS row stride of the tile accumulators
template <bool ZON, bool FRESH>
__global__ void __launch_bounds__(RBLK) render_tile_kernel(RenderConst c, const int *__restrict__ tcnt, const int *__restrict__ toff,
                                                           const float4 *__restrict__ ent, float *__restrict__ img,
                                                           float *__restrict__ zimg)
{
    __shared__ __align__(16) float acc[RTILE * RSTR];
    __shared__ __align__(16) float zacc[ZON ? RTILE * RSTR : 1];
    // drawing parameters of up to RBLK listed localizations, one loaded per thread
    __shared__ float4 s_p[RBLK];   // {pp, rs, fx, fy}: one 16-byte broadcast read per drawn entry
    __shared__ float s_z[ZON ? RBLK : 1];
    __shared__ int s_xy[RBLK];
    const int t = blockIdx.x;
    const int cnt = tcnt[t], first = toff[t];
    if (!FRESH && cnt == 0) return;   // uniform over the work-group
    const int x0 = (t % c.tiles_x) * RTILE, y0 = (t / c.tiles_x) * RTILE;
    const bool aligned = ((reinterpret_cast<uintptr_t>(img) | (ZON ? reinterpret_cast<uintptr_t>(zimg) : 0)) & 15) == 0;
    const bool vec = aligned && (c.srw & 3) == 0 && x0 + RTILE <= c.srw && y0 + RTILE <= c.srh;   // uniform
    if (FRESH && cnt == 0) {   // an empty tile of a fresh image
        if (vec) {
#pragma unroll
            for (int j = 0; j < RVEC; j++) {
                const int p = 4 * (threadIdx.x + j * RBLK);
                const size_t g = (size_t)(y0 + p / RTILE) * c.srw + (x0 + p % RTILE);
                *reinterpret_cast<float4 *>(img + g) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (ZON) *reinterpret_cast<float4 *>(zimg + g) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
            return;
        }
        for (int p = threadIdx.x; p < RTILE * RTILE; p += RBLK) {
            const int gy = y0 + p / RTILE, gx = x0 + p % RTILE;
            if (gx >= c.srw || gy >= c.srh) continue;
            const size_t g = (size_t)gy * c.srw + gx;
            img[g] = 0.0f;
            if (ZON) zimg[g] = 0.0f;
        }
        return;
    }
    float4 v[RVEC], u[RVEC];
    if (vec && !FRESH) {
#pragma unroll
        for (int j = 0; j < RVEC; j++) {
            const int p = 4 * (threadIdx.x + j * RBLK);
            const size_t g = (size_t)(y0 + p / RTILE) * c.srw + (x0 + p % RTILE);
            v[j] = *reinterpret_cast<const float4 *>(img + g);
            if (ZON) u[j] = *reinterpret_cast<const float4 *>(zimg + g);
        }
    }
    for (int p = threadIdx.x; p < RTILE * RTILE; p += RBLK) {
        acc[(p / RTILE) * RSTR + p % RTILE] = 0.0f;
        if (ZON) zacc[(p / RTILE) * RSTR + p % RTILE] = 0.0f;
    }
    __syncthreads();
    const float4 *te = ent + (size_t)first * RENT;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int a = lane / 7, b = lane - 7 * (lane / 7);   // footprint pixel of this lane (lanes < 49)
    const bool one = !c.is3d && c.mode == QCS_RENDER_PEAKPHOTON_1PIXEL;   // single-pixel splats
    // Rounds of RBLK localizations: every thread stages one entry (one round of memory latency for
    // the whole round, coalesced), then each wave draws every fourth staged localization.
    for (int base = 0; base < cnt; base += RBLK) {
        const int m = min(RBLK, cnt - base);
        if ((int)threadIdx.x < m) {
            const float4 e0 = te[(size_t)(base + threadIdx.x) * RENT], e1 = te[(size_t)(base + threadIdx.x) * RENT + 1];
            s_p[threadIdx.x] = e0;
            if (ZON) s_z[threadIdx.x] = e1.x;
            s_xy[threadIdx.x] = __float_as_int(e1.y);
        }
        __syncthreads();
        for (int k = wave; k < m; k += RBLK / 64) {
            const float4 sp = s_p[k];
            const float pp = sp.x;
            const int xy = s_xy[k];
            const int rx = (int)(short)(xy & 0xffff), ry = xy >> 16;   // tile-relative centre pixel
            if (one) {
                if (lane == 0) atomicAdd(&acc[ry * RSTR + rx], pp);
                continue;
            }
            const int py = ry - 3 + a, px = rx - 3 + b;
            if (lane < 49 && py >= 0 && py < RTILE && px >= 0 && px < RTILE) {
                const float dy = (float)a - sp.w, dx = (float)b - sp.z;
                const float psf = qexpf(__builtin_fmaf(dx, dx, dy * dy) * sp.y);
                atomicAdd(&acc[py * RSTR + px], pp * psf);
                if (ZON) atomicAdd(&zacc[py * RSTR + px], (s_z[k] * pp) * psf);
            }
        }
        __syncthreads();   // the next round overwrites the staged parameters; acc is complete after the last (cnt >= 1 here)
    }
    // the tile is owned by this work-group: plain coalesced writes
    if (vec) {
#pragma unroll
        for (int j = 0; j < RVEC; j++) {
            const int p = 4 * (threadIdx.x + j * RBLK);
            const size_t g = (size_t)(y0 + p / RTILE) * c.srw + (x0 + p % RTILE);
            const float4 ac = *reinterpret_cast<const float4 *>(&acc[(p / RTILE) * RSTR + p % RTILE]);
            if (FRESH) {   // 0 + a == a: the same bits as clearing first
                v[j] = ac;
            } else {
                v[j].x += ac.x;
                v[j].y += ac.y;
                v[j].z += ac.z;
                v[j].w += ac.w;
            }
            *reinterpret_cast<float4 *>(img + g) = v[j];
            if (ZON) {
                const float4 zc = *reinterpret_cast<const float4 *>(&zacc[(p / RTILE) * RSTR + p % RTILE]);
                if (FRESH) {
                    u[j] = zc;
                } else {
                    u[j].x += zc.x;
                    u[j].y += zc.y;
                    u[j].z += zc.z;
                    u[j].w += zc.w;
                }
                *reinterpret_cast<float4 *>(zimg + g) = u[j];
            }
        }
        return;
    }
    for (int p = threadIdx.x; p < RTILE * RTILE; p += RBLK) {
        const int gy = y0 + p / RTILE, gx = x0 + p % RTILE;
        if (gx >= c.srw || gy >= c.srh) continue;
        const size_t g = (size_t)gy * c.srw + gx;
        const int a2 = (p / RTILE) * RSTR + p % RTILE;
        img[g] = FRESH ? acc[a2] : img[g] + acc[a2];
        if (ZON) zimg[g] = FRESH ? zacc[a2] : zimg[g] + zacc[a2];
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

int render_launch(qcs_ctx *ctx, const float *d_loc, int n, float *d_img, float *d_zimg, bool fresh, hipStream_t s)
{
    const qcs_params &p = ctx->p;
    if (n <= 0) {
        if (fresh) {   // nothing to draw: the planes are cleared (a fill kernel, graph-safe)
            const size_t npx = (size_t)p.SRImageWidth * p.SRImageHigh;
            fill_async(reinterpret_cast<int *>(d_img), npx, 0, s);
            if (d_zimg && (p.LocType == QCS_LOCTYPE_AS3D || p.LocType == QCS_LOCTYPE_DH3D))
                fill_async(reinterpret_cast<int *>(d_zimg), npx, 0, s);
            QCS_HIP(hipGetLastError());
        }
        return QCS_OK;
    }
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
    c.tiles_x = (c.srw + RTILE - 1) / RTILE;
    c.tiles_y = (c.srh + RTILE - 1) / RTILE;
    const size_t ntile = (size_t)c.tiles_x * c.tiles_y;
    if (ntile > (size_t)INT32_MAX / 8 || (size_t)n * 4 > (size_t)INT32_MAX) return fail(QCS_ERR_ARG, "SR image or row count too large");
    const int nst = scan_tile_count(ntile);
    // tile counts, offsets (+ scan sums), cursors, then the entries (16-byte aligned; a footprint
    // touches <= 4 tiles)
    const size_t head = (ntile * 3 + (size_t)nst + 1 + 3) & ~(size_t)3;
    const size_t words = head + (size_t)n * 4 * RENT * 4;
    int rc;
    if ((rc = ctx->rn_buf.ensure(words * sizeof(int)))) return rc;
    int *tcnt = ctx->rn_buf.as<int>(), *toff = tcnt + ntile, *tsum = toff + ntile, *cursor = tsum + nst + 1;
    float4 *ent = reinterpret_cast<float4 *>(ctx->rn_buf.as<int>() + head);
    fill_async(tcnt, ntile, 0, s);
    fill_async(cursor, ntile, 0, s);
    const dim3 g1((n + 255) / 256), b1(256);
    hipLaunchKernelGGL(render_bin_kernel, g1, b1, 0, s, d_loc, n, c, tcnt, nullptr, nullptr, nullptr);
    if ((rc = scan_exclusive(tcnt, (int)ntile, toff, tsum, s))) return rc;
    hipLaunchKernelGGL(render_bin_kernel, g1, b1, 0, s, d_loc, n, c, tcnt, toff, cursor, ent);
    const dim3 tg((unsigned)ntile), tb(RBLK);
    if (c.is3d && d_zimg) {
        if (fresh) hipLaunchKernelGGL((render_tile_kernel<true, true>), tg, tb, 0, s, c, tcnt, toff, ent, d_img, d_zimg);
        else hipLaunchKernelGGL((render_tile_kernel<true, false>), tg, tb, 0, s, c, tcnt, toff, ent, d_img, d_zimg);
    } else {
        if (fresh) hipLaunchKernelGGL((render_tile_kernel<false, true>), tg, tb, 0, s, c, tcnt, toff, ent, d_img, nullptr);
        else hipLaunchKernelGGL((render_tile_kernel<false, false>), tg, tb, 0, s, c, tcnt, toff, ent, d_img, nullptr);
    }
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
    // flag, offsets, tile sums, keys in/out, indices; sort workspace
    const size_t words = (size_t)n * 5 + ntiles + 1;
    int rc;
    if ((rc = c->fit_flag.ensure(sizeof(int) * words)) || (rc = c->fit_rows.ensure(sizeof(int) * radix_sort_tmp_ints(n))) ||
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
        unsigned *kres;
        int *ires;
        if ((rc = radix_sort_pairs_u32(kin, iin, kout, iout, m, bits, c->fit_rows.as<int>(), s, &kres, &ires))) return rc;
        fill_async(c->fit_cnt.as<int>(), 1, m, s);
        hipLaunchKernelGGL(copy_rows_kernel, dim3((m + 255) / 256), dim3(256), 0, s, d_loc, ires, c->fit_cnt.as<int>(), d_out);
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
    rc = render_launch(c, d_loc, n, d_img, d_zimg, false, (hipStream_t)stream);
    if (rc) return rc;
    return timer_end(c->timer, TM_RENDER, (hipStream_t)stream, ev, n);
}

extern "C" int qcs_render_fresh_dev(qcs_ctx *c, const float *d_loc, int n, float *d_img, float *d_zimg, void *stream)
{
    if (!c || n < 0 || !d_img || (n > 0 && !d_loc)) return fail(QCS_ERR_ARG, "invalid argument");
    if (c->p.SRImageWidth <= 16 || c->p.SRImageHigh <= 16) return fail(QCS_ERR_ARG, "SR image size not set");
    hipEvent_t ev;
    int rc = timer_begin(c->timer, (hipStream_t)stream, &ev);
    if (rc) return rc;
    rc = render_launch(c, d_loc, n, d_img, d_zimg, true, (hipStream_t)stream);
    if (rc) return rc;
    return timer_end(c->timer, TM_RENDER, (hipStream_t)stream, ev, n);
}
