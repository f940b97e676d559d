// nn.hip -- the learned warm start: a small CNN + MLP trained and run on gfx950 with BF16 MFMA
// GEMMs (FP32 accumulate, FP32 master weights, Adam).  No counterpart in the reference.
//
// Network (7 x 7 ROI, valid convolutions, ReLU after every layer but the last):
//   conv1 1 -> 32 (3x3) : 5x5x32      conv2 32 -> 64 : 3x3x64      conv3 64 -> 64 : 1x1x64
//   concat [64 features, ln(scale) / 10, bhat / scale] (zero-padded to 80)
//   fc1 80 -> 64        fc2 64 -> 16 : [3 count logits, (x/R, y/R, N/(2 pi scale)) x 3 slots,
//                                       s - 1.3, (b - bhat) / scale, 0, 0]
// Input: q = max((ADU - Offset) KAdc, 1) / QE, bhat = the edge background estimate,
// scale = max(q) - bhat + 1, x = (q - bhat) / scale.
//
// Every layer is an "NT" GEMM (nn_gemm.h) over im2col rows: conv1 has 25 rows per ROI of 9 taps
// padded to 16; conv2 9 rows of 288; conv3 reads conv2's output rows as one 576-wide row per ROI
// (no copy).  Backward: dW = dY^T X and dX = dY W through transposed bf16 copies, col2im for
// conv2 -> conv1, bias gradients by column sums.
#include <math.h>
#include <string.h>
#include <new>
#include <vector>
#include "nn_gemm.h"
#include "qcs_internal.h"

namespace qcs {

namespace nnk {

constexpr int NR = 7, NRR = 49;
constexpr int C1 = 32, C2 = 64, C3 = 64, F0 = 80, F1 = 64, NO = 16, NSIDE = F0 - C3;
constexpr int K1 = 16;        // conv1 taps (9) padded
constexpr int K2 = 9 * C1;    // 288
constexpr int K3 = 9 * C2;    // 576
// parameter blocks: fp32 master layout == bf16 GEMM layout, rows = output channels
constexpr int W1N = C1 * K1, B1N = C1, W2N = C2 * K2, B2N = C2, W3N = C3 * K3, B3N = C3, W4N = F1 * F0, B4N = F1,
              W5N = NO * F1, B5N = NO;
constexpr int OW1 = 0, OB1 = OW1 + W1N, OW2 = OB1 + B1N, OB2 = OW2 + W2N, OW3 = OB2 + B2N, OB3 = OW3 + W3N,
              OW4 = OB3 + B3N, OB4 = OW4 + W4N, OW5 = OB4 + B4N, OB5 = OW5 + W5N, NPARAM = OB5 + B5N;
constexpr int CHUNK = 65536;  // ROIs per forward/backward pass (activation buffers scale with it)
constexpr float CE_W = 4.0f;  // weight of the count cross-entropy against the parameter errors

inline int rup16(int v) { return (v + 15) & ~15; }
inline unsigned nblk(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// per ROI: normalised image -> conv1 im2col rows, side features, (scale, bhat)
__global__ void prep_kernel(const uint16_t *__restrict__ rois, int n, float offset, float kadc, float inv_qe,
                            uint16_t *__restrict__ x1, float *__restrict__ norm, uint16_t *__restrict__ side)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint16_t *src = rois + (size_t)i * NR * (NR + 1);
    float q[NRR];
    float mx = 0.0f;
#pragma unroll
    for (int k = 0; k < NRR; k++) {
        q[k] = fmaxf(((float)src[k] - offset) * kadc, 1.0f) * inv_qe;
        mx = fmaxf(mx, q[k]);
    }
    float e0 = 0, e1 = 0, e2 = 0, e3 = 0;
#pragma unroll
    for (int k = 0; k < NR; k++) {
        e0 += q[k];
        e1 += q[(NR - 1) * NR + k];
        e2 += q[k * NR];
        e3 += q[k * NR + NR - 1];
    }
    const float bh = 0.5f * (fminf(e0, e1) + fminf(e2, e3)) / (float)NR;
    const float scale = fmaxf(mx - bh, 0.0f) + 1.0f, is = 1.0f / scale;
    norm[2 * i] = scale;
    norm[2 * i + 1] = bh;
    uint16_t xv[NRR];
#pragma unroll
    for (int k = 0; k < NRR; k++) xv[k] = f2bf((q[k] - bh) * is);
    uint16_t *o = x1 + (size_t)i * 25 * K1;
#pragma unroll
    for (int py = 0; py < 5; py++)
#pragma unroll
        for (int px = 0; px < 5; px++) {
            uint16_t *row = o + (py * 5 + px) * K1;
#pragma unroll
            for (int t = 0; t < K1; t++) row[t] = (t < 9) ? xv[(py + t / 3) * NR + px + t % 3] : (uint16_t)0;
        }
    uint16_t *sd = side + (size_t)i * NSIDE;
    sd[0] = f2bf(logf(scale) * 0.1f);
    sd[1] = f2bf(bh * is);
#pragma unroll
    for (int t = 2; t < NSIDE; t++) sd[t] = 0;
}

// out(bf16, ldo) = act(Y + bias)
__global__ void bias_act_kernel(const float *__restrict__ Y, const float *__restrict__ bias, int rows, int ncol, int ldy,
                                uint16_t *__restrict__ out, int ldo)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * ncol) return;
    const int r = (int)(i / ncol), c = (int)(i % ncol);
    out[(size_t)r * ldo + c] = f2bf(fmaxf(Y[(size_t)r * ldy + c] + bias[c], 0.0f));
}

__global__ void out_bias_kernel(float *__restrict__ yo, const float *__restrict__ b, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * NO) yo[i] += b[i % NO];
}

__global__ void side_kernel(const uint16_t *__restrict__ side, int n, uint16_t *__restrict__ a3)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * NSIDE) return;
    a3[(size_t)(i / NSIDE) * F0 + C3 + i % NSIDE] = side[i];
}

// conv2 im2col: a1 [n*25][32] -> x2 [n*9][288], column (ty*3 + tx)*32 + ch; 4 channels per thread
__global__ void im2col2_kernel(const uint16_t *__restrict__ a1, int n, uint16_t *__restrict__ x2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * 81 * (C1 / 4)) return;
    const int c4 = (int)(i % (C1 / 4));
    size_t t = i / (C1 / 4);
    const int tap = (int)(t % 9);
    t /= 9;
    const int pos = (int)(t % 9);
    const size_t roi = t / 9;
    const int py = pos / 3, px = pos % 3, ty = tap / 3, tx = tap % 3;
    const uint2 v = *reinterpret_cast<const uint2 *>(a1 + ((roi * 25 + (py + ty) * 5 + (px + tx)) * C1 + c4 * 4));
    *reinterpret_cast<uint2 *>(x2 + ((roi * 9 + pos) * K2 + tap * C1 + c4 * 4)) = v;
}

// col2im for conv2: dx2 [n*9][288] -> da1 [n*25][32] (sum over the taps that read each input)
__global__ void col2im2_kernel(const float *__restrict__ dx2, int n, float *__restrict__ da1)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * 25 * C1) return;
    const int ch = (int)(i % C1);
    const size_t t = i / C1;
    const int p1 = (int)(t % 25);
    const size_t roi = t / 25;
    const int y1 = p1 / 5, x1 = p1 % 5;
    float s = 0.0f;
    for (int ty = 0; ty < 3; ty++)
        for (int tx = 0; tx < 3; tx++) {
            const int py = y1 - ty, px = x1 - tx;
            if (py < 0 || py >= 3 || px < 0 || px >= 3) continue;
            s += dx2[(roi * 9 + py * 3 + px) * K2 + (ty * 3 + tx) * C1 + ch];
        }
    da1[i] = s;
}

// d = dA (x) [a > 0]: fp32 (dense rows x ncol) and bf16 copies
__global__ void relu_back_kernel(const float *__restrict__ dA, int lda, const uint16_t *__restrict__ a, int ld_a, int rows,
                                 int ncol, float *__restrict__ d, uint16_t *__restrict__ db)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * ncol) return;
    const int r = (int)(i / ncol), c = (int)(i % ncol);
    const float v = (bf2f(a[(size_t)r * ld_a + c]) > 0.0f) ? dA[(size_t)r * lda + c] : 0.0f;
    d[i] = v;
    db[i] = f2bf(v);
}

// bf16 transpose with zero padding: in [rows][cols] (ldi) -> out [colsp][rowsp]
__global__ void transpose_kernel(const uint16_t *__restrict__ in, int rows, int cols, int ldi, uint16_t *__restrict__ out,
                                 int rowsp, int colsp)
{
    __shared__ uint16_t tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 256 threads: 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int r = by + k, c = bx + tx;
        tile[k][tx] = (r < rows && c < cols) ? in[(size_t)r * ldi + c] : (uint16_t)0;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = bx + k, r = by + tx;   // out[c][r] = in[r][c]
        if (c < colsp && r < rowsp) out[(size_t)c * rowsp + r] = tile[tx][k];
    }
}

// out[c] += sum over a block of rows (out zeroed beforehand); grid.y splits the rows
__global__ void colsum_kernel(const float *__restrict__ Y, int rows, int ncol, float *__restrict__ out)
{
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int part = blockIdx.y * (blockDim.x >> 6) + (threadIdx.x >> 6), nparts = gridDim.y * (blockDim.x >> 6);
    if (c >= ncol) return;
    float s = 0.0f;
    for (int r = part; r < rows; r += nparts) s += Y[(size_t)r * ncol + c];
    atomicAdd(&out[c], s);
}

__global__ void f2bf_kernel(const float *__restrict__ in, size_t n, uint16_t *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

__global__ void fill0_kernel(float *__restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// Loss and gradient per ROI.  labels [n][16]: K (0 = ignore), (x, y, N) x 3 (ROI pixels,
// photons; slots sorted by N descending), s, b.  Loss = 4 CE(count) + R^2 |dpos/R|^2 per slot
// + (dN / (2 pi scale))^2 per slot + ds^2 + 10 (db / scale)^2.  Gradients are of the sum over the ROIs times
// `wscale` (1 / total labelled ROIs of the step).
__global__ void loss_kernel(const float *__restrict__ out, const float *__restrict__ labels, const float *__restrict__ norm,
                            int n, float wscale, float *__restrict__ dY, float *__restrict__ loss_sum)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float L = 0.0f;
    if (i < n) {
        const float *o = out + (size_t)i * NO, *lb = labels + (size_t)i * NO;
        float *d = dY + (size_t)i * NO;
        float g[NO];
#pragma unroll
        for (int t = 0; t < NO; t++) g[t] = 0.0f;
        const int K = (int)lb[0];
        if (K >= 1 && K <= 3) {
            const float scale = norm[2 * i], bh = norm[2 * i + 1];
            const float m = fmaxf(o[0], fmaxf(o[1], o[2]));
            const float e0 = expf(o[0] - m), e1 = expf(o[1] - m), e2 = expf(o[2] - m), es = e0 + e1 + e2;
            const float p[3] = {e0 / es, e1 / es, e2 / es};
            L += -CE_W * logf(fmaxf(p[K - 1], 1e-30f));
#pragma unroll
            for (int t = 0; t < 3; t++) g[t] = CE_W * (p[t] - (t == K - 1 ? 1.0f : 0.0f));
            const float itgt = 1.0f / (6.283185307179586f * scale);
            const float wp = (float)(NR * NR);
            for (int k = 0; k < K; k++) {
                const float tx = lb[1 + 3 * k] / (float)NR, ty = lb[2 + 3 * k] / (float)NR, tn = lb[3 + 3 * k] * itgt;
                const float ex = o[3 + 3 * k] - tx, ey = o[4 + 3 * k] - ty, en = o[5 + 3 * k] - tn;
                L += wp * (ex * ex + ey * ey) + en * en;
                g[3 + 3 * k] = 2.0f * wp * ex;
                g[4 + 3 * k] = 2.0f * wp * ey;
                g[5 + 3 * k] = 2.0f * en;
            }
            const float es2 = o[12] - (lb[10] - 1.3f), eb = o[13] - (lb[11] - bh) / scale;
            L += es2 * es2 + 10.0f * eb * eb;
            g[12] = 2.0f * es2;
            g[13] = 20.0f * eb;
        }
#pragma unroll
        for (int t = 0; t < NO; t++) d[t] = g[t] * wscale;
    }
    for (int off = 32; off > 0; off >>= 1) L += __shfl_down(L, off, 64);
    if ((threadIdx.x & 63) == 0 && L != 0.0f) atomicAdd(loss_sum, L);
}

__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                            int n, float lr, float b1, float b2, float c1, float c2, float eps)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    if (i >= OW1 && i < OW1 + W1N && ((i - OW1) % K1) >= 9) gi = 0.0f;   // padded conv1 taps stay zero
    const float mi = fmaf(b1, m[i], (1.0f - b1) * gi), vi = fmaf(b2, v[i], (1.0f - b2) * gi * gi);
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// ---- fused inference ------------------------------------------------------------------------
// One wave computes the whole network for 16 ROIs.  Every layer is evaluated transposed,
// D^T[out channel][roi] = W[out][k] . act^T[k][roi]: the weights are the MFMA A operand (8-byte
// fragments straight from L2) and the activations the B operand.  An MFMA result then already has
// the B-operand layout of the next layer (lane l holds channels 4 (l / 16) .. +3 of ROI l % 16),
// so bias, ReLU and the bf16 rounding are applied in registers and no activation touches LDS or
// HBM.  Per output the k-chunks are accumulated in the order of the layer-by-layer GEMMs, with
// the same bf16 rounding points.
__device__ __forceinline__ bf16x4_t act4(const f32x4_t &d, const float *bias)
{
    bf16x4_t v;
#pragma unroll
    for (int i = 0; i < 4; i++) v[i] = (short)f2bf(fmaxf(d[i] + bias[i], 0.0f));
    return v;
}

__device__ __forceinline__ bf16x4_t wfrag(const uint16_t *W, int row, int ld, int k)
{
    return *reinterpret_cast<const bf16x4_t *>(W + (size_t)row * ld + k);
}

__device__ __forceinline__ f32x4_t mfma(const bf16x4_t &a, const bf16x4_t &b, const f32x4_t &c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__global__ void __launch_bounds__(64) fwd_fused_kernel(const uint16_t *__restrict__ rois, int n, float offset, float kadc,
                                                       float inv_qe, const uint16_t *__restrict__ W, const float *__restrict__ wf,
                                                       float *__restrict__ out, float *__restrict__ norm)
{
    __shared__ uint16_t xs[16][52];   // normalised input, bf16 (49 used)
    __shared__ uint16_t sides[16][2];
    __shared__ bf16x4_t c1s[3][5][2][64];   // conv1 tiles of three input rows, in B-operand layout
    const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
    const int base = blockIdx.x * 16;
    if (lane < 16) {   // input normalisation of ROI base + lane (as prep_kernel)
        const int i = base + lane;
        if (i < n) {
            const uint16_t *src = rois + (size_t)i * NR * (NR + 1);
            float q[NRR];
            float mx = 0.0f;
#pragma unroll
            for (int k = 0; k < NRR; k++) {
                q[k] = fmaxf(((float)src[k] - offset) * kadc, 1.0f) * inv_qe;
                mx = fmaxf(mx, q[k]);
            }
            float e0 = 0, e1 = 0, e2 = 0, e3 = 0;
#pragma unroll
            for (int k = 0; k < NR; k++) {
                e0 += q[k];
                e1 += q[(NR - 1) * NR + k];
                e2 += q[k * NR];
                e3 += q[k * NR + NR - 1];
            }
            const float bh = 0.5f * (fminf(e0, e1) + fminf(e2, e3)) / (float)NR;
            const float scale = fmaxf(mx - bh, 0.0f) + 1.0f, is = 1.0f / scale;
            if (norm) {
                norm[2 * i] = scale;
                norm[2 * i + 1] = bh;
            }
#pragma unroll
            for (int k = 0; k < NRR; k++) xs[lane][k] = f2bf((q[k] - bh) * is);
            sides[lane][0] = f2bf(logf(scale) * 0.1f);
            sides[lane][1] = f2bf(bh * is);
        } else {
#pragma unroll
            for (int k = 0; k < NRR; k++) xs[lane][k] = 0;
            sides[lane][0] = sides[lane][1] = 0;
        }
    }
    __syncthreads();
    const float *b1 = wf + OB1, *b2 = wf + OB2, *b3 = wf + OB3, *b4 = wf + OB4, *b5 = wf + OB5;
    const f32x4_t zero = {0.0f, 0.0f, 0.0f, 0.0f};
    bf16x4_t w1[2];
#pragma unroll
    for (int t = 0; t < 2; t++) w1[t] = wfrag(W + OW1, t * 16 + r, K1, 4 * g);
    f32x4_t acc3[4] = {zero, zero, zero, zero};
#pragma unroll 1
    for (int py = 0; py < 3; py++) {
        // conv1 at input rows py .. py + 2 (all 5 columns, both 16-channel tiles); each lane
        // reads back only what it wrote, so no barrier is needed
#pragma unroll
        for (int yy = 0; yy < 3; yy++)
#pragma unroll
            for (int xx = 0; xx < 5; xx++) {
                bf16x4_t xb;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int t = 4 * g + e;   // tap; 9..15 are padding
                    const int ty = t / 3, tx = t % 3;
                    xb[e] = t < 9 ? (short)xs[r][(py + yy + ty) * NR + xx + tx] : (short)0;
                }
#pragma unroll
                for (int ct = 0; ct < 2; ct++) c1s[yy][xx][ct][lane] = act4(mfma(w1[ct], xb, zero), b1 + ct * 16 + 4 * g);
            }
        // conv2 at the three outputs of row py; k = tap * 32 + channel, ascending
        f32x4_t acc2[3][4];
#pragma unroll
        for (int px = 0; px < 3; px++)
#pragma unroll
            for (int ot = 0; ot < 4; ot++) acc2[px][ot] = zero;
#pragma unroll 1
        for (int tap = 0; tap < 9; tap++) {
            const int ty = tap / 3, tx = tap % 3;
#pragma unroll
            for (int ct = 0; ct < 2; ct++) {
                const int k0 = tap * C1 + ct * 16 + 4 * g;
                bf16x4_t wa[4];
#pragma unroll
                for (int ot = 0; ot < 4; ot++) wa[ot] = wfrag(W + OW2, ot * 16 + r, K2, k0);
#pragma unroll
                for (int px = 0; px < 3; px++)
#pragma unroll
                    for (int ot = 0; ot < 4; ot++) acc2[px][ot] = mfma(wa[ot], c1s[ty][px + tx][ct][lane], acc2[px][ot]);
            }
        }
        // conv3 over these positions; k = position * 64 + channel, ascending
#pragma unroll
        for (int px = 0; px < 3; px++) {
            const int p2 = py * 3 + px;
#pragma unroll
            for (int ct = 0; ct < 4; ct++) {
                const bf16x4_t a = act4(acc2[px][ct], b2 + ct * 16 + 4 * g);
                const int k0 = p2 * C2 + ct * 16 + 4 * g;
#pragma unroll
                for (int ot = 0; ot < 4; ot++) acc3[ot] = mfma(wfrag(W + OW3, ot * 16 + r, K3, k0), a, acc3[ot]);
            }
        }
    }
    // fc1 over [conv3 features, side features, zero padding]
    f32x4_t acc4[4] = {zero, zero, zero, zero};
#pragma unroll
    for (int ct = 0; ct < 4; ct++) {
        const bf16x4_t a = act4(acc3[ct], b3 + ct * 16 + 4 * g);
#pragma unroll
        for (int ot = 0; ot < 4; ot++) acc4[ot] = mfma(wfrag(W + OW4, ot * 16 + r, F0, ct * 16 + 4 * g), a, acc4[ot]);
    }
    {
        bf16x4_t sb = {0, 0, 0, 0};
        if (g == 0) {
            sb[0] = (short)sides[r][0];
            sb[1] = (short)sides[r][1];
        }
#pragma unroll
        for (int ot = 0; ot < 4; ot++) acc4[ot] = mfma(wfrag(W + OW4, ot * 16 + r, F0, C3 + 4 * g), sb, acc4[ot]);
    }
    // fc2
    f32x4_t acc5 = zero;
#pragma unroll
    for (int ct = 0; ct < 4; ct++) acc5 = mfma(wfrag(W + OW5, r, F1, ct * 16 + 4 * g), act4(acc4[ct], b4 + ct * 16 + 4 * g), acc5);
    const int i = base + r;
    if (i < n) {
        float4 o;
        o.x = acc5[0] + b5[4 * g + 0];
        o.y = acc5[1] + b5[4 * g + 1];
        o.z = acc5[2] + b5[4 * g + 2];
        o.w = acc5[3] + b5[4 * g + 3];
        *reinterpret_cast<float4 *>(out + (size_t)i * NO + 4 * g) = o;
    }
}

}  // namespace nnk

void nn_gemm_nt(const uint16_t *A, const uint16_t *B, float *C, int M, int N, int K, int lda, int ldb, int ldc, bool acc,
                hipStream_t s)
{
    if (M <= 0 || N <= 0) return;
    const dim3 grid((N + 63) / 64, (M + 63) / 64), block(256);
    if (acc) hipLaunchKernelGGL(gemm_nt_kernel<true>, grid, block, 0, s, A, B, C, M, N, K, lda, ldb, ldc);
    else hipLaunchKernelGGL(gemm_nt_kernel<false>, grid, block, 0, s, A, B, C, M, N, K, lda, ldb, ldc);
}

}  // namespace qcs

using namespace qcs;
using namespace qcs::nnk;

struct qcs_net {
    int device = 0;
    long long steps = 0;
    DevBuf w, g, m, v, wb, wt, loss;          // parameters (fp32 master, gradient, Adam moments), bf16 copy
    DevBuf x1, a1, x2, a2, a3, a4, side, norm, yo, Y, dF, dG, dB, tA, tB;
    HostBuf pin;
    bool ready = false;
};

namespace {

int net_alloc(qcs_net *nt)
{
    const size_t N = CHUNK;
    int rc;
    if ((rc = nt->w.ensure(NPARAM * 4)) || (rc = nt->g.ensure(NPARAM * 4)) || (rc = nt->m.ensure(NPARAM * 4)) ||
        (rc = nt->v.ensure(NPARAM * 4)) || (rc = nt->wb.ensure(NPARAM * 2)) || (rc = nt->wt.ensure((size_t)K3 * C3 * 2)) ||
        (rc = nt->loss.ensure(64)) || (rc = nt->x1.ensure(N * 25 * K1 * 2)) || (rc = nt->a1.ensure(N * 25 * C1 * 2)) ||
        (rc = nt->x2.ensure(N * 9 * K2 * 2)) || (rc = nt->a2.ensure(N * 9 * C2 * 2)) || (rc = nt->a3.ensure(N * F0 * 2)) ||
        (rc = nt->a4.ensure(N * F1 * 2)) || (rc = nt->side.ensure(N * NSIDE * 2)) || (rc = nt->norm.ensure(N * 2 * 4)) ||
        (rc = nt->yo.ensure(N * NO * 4)) || (rc = nt->Y.ensure(N * 25 * C1 * 4)) || (rc = nt->dF.ensure(N * 9 * K2 * 4)) ||
        (rc = nt->dG.ensure(N * 25 * C1 * 4)) || (rc = nt->dB.ensure(N * 25 * C1 * 2)) ||
        (rc = nt->tA.ensure(N * 25 * C1 * 2)) || (rc = nt->tB.ensure(N * 9 * K2 * 2)) || (rc = nt->pin.ensure(64)))
        return rc;
    return QCS_OK;
}

void tr(const uint16_t *in, int rows, int cols, int ldi, uint16_t *out, hipStream_t s)
{
    const int rp = rup16(rows), cp = rup16(cols);
    hipLaunchKernelGGL(transpose_kernel, dim3((cp + 31) / 32, (rp + 31) / 32), dim3(256), 0, s, in, rows, cols, ldi, out, rp, cp);
}

void colsum(const float *Y, int rows, int ncol, float *out, hipStream_t s)
{
    int parts = (rows + 4095) / 4096;
    parts = parts < 1 ? 1 : (parts > 64 ? 64 : parts);
    hipLaunchKernelGGL(colsum_kernel, dim3((ncol + 63) / 64, parts), dim3(256), 0, s, Y, rows, ncol, out);
}

void refresh_bf16(qcs_net *nt, hipStream_t s)
{
    hipLaunchKernelGGL(f2bf_kernel, dim3(nblk(NPARAM)), dim3(256), 0, s, nt->w.as<float>(), (size_t)NPARAM, nt->wb.as<uint16_t>());
}

// forward over n <= CHUNK ROIs: activations, outputs yo [n][16] and norm [n][2]
int net_forward(qcs_net *nt, const qcs_params &p, const uint16_t *d_rois, int n, hipStream_t s)
{
    const uint16_t *W = nt->wb.as<uint16_t>();
    const float *wf = nt->w.as<float>();
    float *Y = nt->Y.as<float>();
    hipLaunchKernelGGL(prep_kernel, dim3(nblk(n, 64)), dim3(64), 0, s, d_rois, n, p.Offset, p.KAdc, 1.0f / p.QE,
                       nt->x1.as<uint16_t>(), nt->norm.as<float>(), nt->side.as<uint16_t>());
    nn_gemm_nt(nt->x1.as<uint16_t>(), W + OW1, Y, n * 25, C1, K1, K1, K1, C1, false, s);
    hipLaunchKernelGGL(bias_act_kernel, dim3(nblk((size_t)n * 25 * C1)), dim3(256), 0, s, Y, wf + OB1, n * 25, C1, C1,
                       nt->a1.as<uint16_t>(), C1);
    hipLaunchKernelGGL(im2col2_kernel, dim3(nblk((size_t)n * 81 * (C1 / 4))), dim3(256), 0, s, nt->a1.as<uint16_t>(), n,
                       nt->x2.as<uint16_t>());
    nn_gemm_nt(nt->x2.as<uint16_t>(), W + OW2, Y, n * 9, C2, K2, K2, K2, C2, false, s);
    hipLaunchKernelGGL(bias_act_kernel, dim3(nblk((size_t)n * 9 * C2)), dim3(256), 0, s, Y, wf + OB2, n * 9, C2, C2,
                       nt->a2.as<uint16_t>(), C2);
    nn_gemm_nt(nt->a2.as<uint16_t>(), W + OW3, Y, n, C3, K3, K3, K3, C3, false, s);
    hipLaunchKernelGGL(bias_act_kernel, dim3(nblk((size_t)n * C3)), dim3(256), 0, s, Y, wf + OB3, n, C3, C3,
                       nt->a3.as<uint16_t>(), F0);
    hipLaunchKernelGGL(side_kernel, dim3(nblk((size_t)n * NSIDE)), dim3(256), 0, s, nt->side.as<uint16_t>(), n,
                       nt->a3.as<uint16_t>());
    nn_gemm_nt(nt->a3.as<uint16_t>(), W + OW4, Y, n, F1, F0, F0, F0, F1, false, s);
    hipLaunchKernelGGL(bias_act_kernel, dim3(nblk((size_t)n * F1)), dim3(256), 0, s, Y, wf + OB4, n, F1, F1,
                       nt->a4.as<uint16_t>(), F1);
    float *yo = nt->yo.as<float>();
    nn_gemm_nt(nt->a4.as<uint16_t>(), W + OW5, yo, n, NO, F1, F1, F1, NO, false, s);
    hipLaunchKernelGGL(out_bias_kernel, dim3(nblk((size_t)n * NO)), dim3(256), 0, s, yo, wf + OB5, n);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

// accumulate the gradient of the loss over n <= CHUNK labelled ROIs (after net_forward) into g
int net_backward(qcs_net *nt, const float *d_labels, int n, float wscale, hipStream_t s)
{
    const int n16 = rup16(n), n9 = rup16(9 * n), n25 = rup16(25 * n);
    float *g = nt->g.as<float>(), *dF = nt->dF.as<float>(), *dG = nt->dG.as<float>();
    uint16_t *W = nt->wb.as<uint16_t>(), *dB = nt->dB.as<uint16_t>(), *tA = nt->tA.as<uint16_t>(), *tB = nt->tB.as<uint16_t>(),
             *wt = nt->wt.as<uint16_t>();
    hipLaunchKernelGGL(loss_kernel, dim3(nblk(n, 64)), dim3(64), 0, s, nt->yo.as<float>(), d_labels, nt->norm.as<float>(), n,
                       wscale, dG, nt->loss.as<float>());
    // fc2
    colsum(dG, n, NO, g + OB5, s);
    hipLaunchKernelGGL(f2bf_kernel, dim3(nblk((size_t)n * NO)), dim3(256), 0, s, dG, (size_t)n * NO, dB);
    tr(dB, n, NO, NO, tA, s);
    tr(nt->a4.as<uint16_t>(), n, F1, F1, tB, s);
    nn_gemm_nt(tA, tB, g + OW5, NO, F1, n16, n16, n16, F1, true, s);
    tr(W + OW5, NO, F1, F1, wt, s);
    nn_gemm_nt(dB, wt, dF, n, F1, NO, NO, NO, F1, false, s);
    hipLaunchKernelGGL(relu_back_kernel, dim3(nblk((size_t)n * F1)), dim3(256), 0, s, dF, F1, nt->a4.as<uint16_t>(), F1, n, F1, dG,
                       dB);
    // fc1
    colsum(dG, n, F1, g + OB4, s);
    tr(dB, n, F1, F1, tA, s);
    tr(nt->a3.as<uint16_t>(), n, F0, F0, tB, s);
    nn_gemm_nt(tA, tB, g + OW4, F1, F0, n16, n16, n16, F0, true, s);
    tr(W + OW4, F1, F0, F0, wt, s);
    nn_gemm_nt(dB, wt, dF, n, F0, F1, F1, F1, F0, false, s);
    hipLaunchKernelGGL(relu_back_kernel, dim3(nblk((size_t)n * C3)), dim3(256), 0, s, dF, F0, nt->a3.as<uint16_t>(), F0, n, C3, dG,
                       dB);
    // conv3 (input: a2 as [n][576])
    colsum(dG, n, C3, g + OB3, s);
    tr(dB, n, C3, C3, tA, s);
    tr(nt->a2.as<uint16_t>(), n, K3, K3, tB, s);
    nn_gemm_nt(tA, tB, g + OW3, C3, K3, n16, n16, n16, K3, true, s);
    tr(W + OW3, C3, K3, K3, wt, s);
    nn_gemm_nt(dB, wt, dF, n, K3, C3, C3, C3, K3, false, s);   // = d a2 [n*9][64]
    hipLaunchKernelGGL(relu_back_kernel, dim3(nblk((size_t)n * 9 * C2)), dim3(256), 0, s, dF, C2, nt->a2.as<uint16_t>(), C2, n * 9,
                       C2, dG, dB);
    // conv2
    colsum(dG, n * 9, C2, g + OB2, s);
    tr(dB, n * 9, C2, C2, tA, s);
    tr(nt->x2.as<uint16_t>(), n * 9, K2, K2, tB, s);
    nn_gemm_nt(tA, tB, g + OW2, C2, K2, n9, n9, n9, K2, true, s);
    tr(W + OW2, C2, K2, K2, wt, s);
    nn_gemm_nt(dB, wt, dF, n * 9, K2, C2, C2, C2, K2, false, s);   // d x2 [n*9][288]
    hipLaunchKernelGGL(col2im2_kernel, dim3(nblk((size_t)n * 25 * C1)), dim3(256), 0, s, dF, n, nt->Y.as<float>());
    hipLaunchKernelGGL(relu_back_kernel, dim3(nblk((size_t)n * 25 * C1)), dim3(256), 0, s, nt->Y.as<float>(), C1,
                       nt->a1.as<uint16_t>(), C1, n * 25, C1, dG, dB);
    // conv1
    colsum(dG, n * 25, C1, g + OB1, s);
    tr(dB, n * 25, C1, C1, tA, s);
    tr(nt->x1.as<uint16_t>(), n * 25, K1, K1, tB, s);
    nn_gemm_nt(tA, tB, g + OW1, C1, K1, n25, n25, n25, K1, true, s);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

// gradient of the mean loss over all n ROIs (chunked) into g; *loss_out = mean loss
int net_gradient(qcs_net *nt, const qcs_params &p, const uint16_t *d_rois, const float *d_labels, int n, float *loss_out,
                 hipStream_t s)
{
    hipLaunchKernelGGL(fill0_kernel, dim3(nblk(NPARAM)), dim3(256), 0, s, nt->g.as<float>(), NPARAM);
    hipLaunchKernelGGL(fill0_kernel, dim3(1), dim3(64), 0, s, nt->loss.as<float>(), 1);
    const float ws = 1.0f / (float)(n > 0 ? n : 1);
    for (int c0 = 0; c0 < n; c0 += CHUNK) {
        const int m = (n - c0) < CHUNK ? (n - c0) : CHUNK;
        int rc = net_forward(nt, p, d_rois + (size_t)c0 * NR * (NR + 1), m, s);
        if (rc) return rc;
        if ((rc = net_backward(nt, d_labels + (size_t)c0 * NO, m, ws, s))) return rc;
    }
    if (loss_out) {
        float *h = nt->pin.as<float>();
        QCS_HIP(hipMemcpyAsync(h, nt->loss.ptr, sizeof(float), hipMemcpyDeviceToHost, s));
        QCS_HIP(hipStreamSynchronize(s));
        *loss_out = h[0] * ws;
    }
    return QCS_OK;
}

// deterministic He-uniform initialisation (splitmix64 stream)
uint64_t splitmix(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void init_params(std::vector<float> &w, uint64_t seed)
{
    w.assign(NPARAM, 0.0f);
    uint64_t st = seed * 0x2545F4914F6CDD1Dull + 1;
    auto fill = [&](int off, int rows, int cols, int fan_in, int used_cols) {
        const float lim = sqrtf(6.0f / (float)fan_in);
        for (int r = 0; r < rows; r++)
            for (int c = 0; c < cols; c++) {
                const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
                w[off + r * cols + c] = (c < used_cols) ? (float)((2.0 * u - 1.0) * lim) : 0.0f;
            }
    };
    fill(OW1, C1, K1, 9, 9);
    fill(OW2, C2, K2, K2, K2);
    fill(OW3, C3, K3, K3, K3);
    fill(OW4, F1, F0, C3 + 2, C3 + 2);
    fill(OW5, NO, F1, F1, F1);
    for (int i = 0; i < F1; i++) w[OW5 + 14 * F1 + i] = w[OW5 + 15 * F1 + i] = 0.0f;   // unused outputs
}

}  // namespace

extern "C" {

int qcs_net_param_count(void) { return NPARAM; }

int qcs_net_create(int roi_size, uint64_t seed, qcs_net **out)
{
    if (!out) return fail(QCS_ERR_ARG, "null argument");
    *out = nullptr;
    if (roi_size != NR) return fail(QCS_ERR_ARG, "the warm-start network takes 7x7 ROIs (got %d)", roi_size);
    qcs_net *nt = new (std::nothrow) qcs_net();
    if (!nt) return fail(QCS_ERR_NOMEM, "network allocation failed");
    int rc;
    if (hipGetDevice(&nt->device) != hipSuccess || (rc = net_alloc(nt))) {
        delete nt;
        return fail(QCS_ERR_NOMEM, "network buffers could not be allocated");
    }
    std::vector<float> w;
    init_params(w, seed);
    if (hipMemcpy(nt->w.ptr, w.data(), NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(nt->m.ptr, 0, NPARAM * 4) != hipSuccess || hipMemset(nt->v.ptr, 0, NPARAM * 4) != hipSuccess) {
        delete nt;
        return fail(QCS_ERR_HIP, "network initialisation failed");
    }
    refresh_bf16(nt, nullptr);
    if (hipDeviceSynchronize() != hipSuccess) {
        delete nt;
        return fail(QCS_ERR_HIP, "network initialisation failed");
    }
    *out = nt;
    return QCS_OK;
}

int qcs_net_destroy(qcs_net *nt)
{
    if (!nt) return QCS_OK;
    DevBuf *b[] = {&nt->w, &nt->g, &nt->m, &nt->v, &nt->wb, &nt->wt, &nt->loss, &nt->x1, &nt->a1, &nt->x2, &nt->a2, &nt->a3,
                   &nt->a4, &nt->side, &nt->norm, &nt->yo, &nt->Y, &nt->dF, &nt->dG, &nt->dB, &nt->tA, &nt->tB};
    for (auto *x : b) x->release();
    nt->pin.release();
    delete nt;
    return QCS_OK;
}

int qcs_net_get_params(qcs_net *nt, float *w)
{
    if (!nt || !w) return fail(QCS_ERR_ARG, "invalid argument");
    QCS_HIP(hipDeviceSynchronize());
    QCS_HIP(hipMemcpy(w, nt->w.ptr, NPARAM * 4, hipMemcpyDeviceToHost));
    return QCS_OK;
}

int qcs_net_set_params(qcs_net *nt, const float *w)
{
    if (!nt || !w) return fail(QCS_ERR_ARG, "invalid argument");
    QCS_HIP(hipDeviceSynchronize());
    QCS_HIP(hipMemcpy(nt->w.ptr, w, NPARAM * 4, hipMemcpyHostToDevice));
    QCS_HIP(hipMemset(nt->m.ptr, 0, NPARAM * 4));
    QCS_HIP(hipMemset(nt->v.ptr, 0, NPARAM * 4));
    nt->steps = 0;
    refresh_bf16(nt, nullptr);
    QCS_HIP(hipDeviceSynchronize());
    return QCS_OK;
}

int qcs_net_forward_dev(qcs_net *nt, const qcs_params *p, const uint16_t *d_rois, int n, float *d_out, float *d_norm,
                        void *stream)
{
    if (!nt || !p || n < 0 || (n > 0 && (!d_rois || !d_out))) return fail(QCS_ERR_ARG, "invalid argument");
    if (p->ROISize != NR) return fail(QCS_ERR_ARG, "the warm-start network takes 7x7 ROIs");
    hipStream_t s = (hipStream_t)stream;
    if (n == 0) return QCS_OK;
    hipLaunchKernelGGL(fwd_fused_kernel, dim3((unsigned)((n + 15) / 16)), dim3(64), 0, s, d_rois, n, p->Offset, p->KAdc, 1.0f / p->QE,
                       nt->wb.as<uint16_t>(), nt->w.as<float>(), d_out, d_norm);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

int qcs_net_forward_layers_dev(qcs_net *nt, const qcs_params *p, const uint16_t *d_rois, int n, float *d_out, float *d_norm,
                               void *stream)
{
    if (!nt || !p || n < 0 || (n > 0 && (!d_rois || !d_out))) return fail(QCS_ERR_ARG, "invalid argument");
    if (p->ROISize != NR) return fail(QCS_ERR_ARG, "the warm-start network takes 7x7 ROIs");
    hipStream_t s = (hipStream_t)stream;
    for (int c0 = 0; c0 < n; c0 += CHUNK) {
        const int m = (n - c0) < CHUNK ? (n - c0) : CHUNK;
        int rc = net_forward(nt, *p, d_rois + (size_t)c0 * NR * (NR + 1), m, s);
        if (rc) return rc;
        QCS_HIP(hipMemcpyAsync(d_out + (size_t)c0 * NO, nt->yo.ptr, (size_t)m * NO * 4, hipMemcpyDeviceToDevice, s));
        if (d_norm) QCS_HIP(hipMemcpyAsync(d_norm + (size_t)c0 * 2, nt->norm.ptr, (size_t)m * 8, hipMemcpyDeviceToDevice, s));
    }
    return QCS_OK;
}

int qcs_net_grad_dev(qcs_net *nt, const qcs_params *p, const uint16_t *d_rois, const float *d_labels, int n, float *h_grad,
                     float *loss_out, void *stream)
{
    if (!nt || !p || n <= 0 || !d_rois || !d_labels) return fail(QCS_ERR_ARG, "invalid argument");
    if (p->ROISize != NR) return fail(QCS_ERR_ARG, "the warm-start network takes 7x7 ROIs");
    hipStream_t s = (hipStream_t)stream;
    int rc = net_gradient(nt, *p, d_rois, d_labels, n, loss_out, s);
    if (rc) return rc;
    if (h_grad) {
        QCS_HIP(hipStreamSynchronize(s));
        QCS_HIP(hipMemcpy(h_grad, nt->g.ptr, NPARAM * 4, hipMemcpyDeviceToHost));
    }
    return QCS_OK;
}

int qcs_net_train_step_dev(qcs_net *nt, const qcs_params *p, const uint16_t *d_rois, const float *d_labels, int n, float lr,
                           float *loss_out, void *stream)
{
    if (!nt || !p || n <= 0 || !d_rois || !d_labels || !(lr > 0.0f)) return fail(QCS_ERR_ARG, "invalid argument");
    if (p->ROISize != NR) return fail(QCS_ERR_ARG, "the warm-start network takes 7x7 ROIs");
    hipStream_t s = (hipStream_t)stream;
    int rc = net_gradient(nt, *p, d_rois, d_labels, n, loss_out, s);
    if (rc) return rc;
    nt->steps++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = (float)(1.0 / (1.0 - pow(b1, (double)nt->steps))), c2 = (float)(1.0 / (1.0 - pow(b2, (double)nt->steps)));
    hipLaunchKernelGGL(adam_kernel, dim3(nblk(NPARAM)), dim3(256), 0, s, nt->w.as<float>(), nt->g.as<float>(), nt->m.as<float>(),
                       nt->v.as<float>(), NPARAM, lr, b1, b2, c1, c2, 1e-8f);
    refresh_bf16(nt, s);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

int qcs_dev_gemm_nt_bf16(const uint16_t *d_a, const uint16_t *d_b, float *d_c, int M, int N, int K, void *stream)
{
    if (!d_a || !d_b || !d_c || M <= 0 || N <= 0 || K <= 0 || (K % 16) != 0) return fail(QCS_ERR_ARG, "invalid argument");
    nn_gemm_nt(d_a, d_b, d_c, M, N, K, K, K, N, false, (hipStream_t)stream);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

}  // extern "C"
