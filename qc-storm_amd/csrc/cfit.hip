// cfit.hip -- converged multi-emitter Poisson-MLE fitting with FP64 model selection and FP64
// full-Fisher CRLB on gfx950.
//
// This is the "plain" path of the batched multi-emitter mode and the labeller of the learned
// warm start; it has no counterpart in the reference (whose fitters run a fixed iteration
// schedule, bfgsMLE_GS2D_core.h:313-426).  Model, pixel (r, c) centred at (c + 0.5, r + 0.5):
//
//   mu = sum_k N_k / (2 pi s^2) exp(-((c + .5 - x_k)^2 + (r + .5 - y_k)^2) / (2 s^2)) + b
//
// with K = 1..3 emitters sharing the PSF width s.  theta = [N_k, x_k, y_k]_k, s, b in photons and
// pixels; the optimiser works on u = theta / scale (N / 1000, b / 100) so the identity is a
// sensible first inverse Hessian.  Loss: Poisson negative log-likelihood sum(mu - q ln mu) with
// q = max((ADU - Offset) * KAdc, 1) / QE, analytic gradient, BFGS with Armijo backtracking and
// box projection, run to convergence (max step < tol) or max_iter.
//
// One ROI per lane; the observed image sits in LDS (lane-interleaved), as does the packed
// symmetric inverse Hessian, leaving VGPRs for the separable model.  The starting inverse Hessian
// is the inverse Fisher information at the initial point, which cuts the iterations about 3x
// against the identity and makes the quasi-Newton step a reliable distance to the optimum: a fit
// has converged when that full step is below tol in every scaled coordinate (or below 20 tol
// where the FP32 loss can no longer resolve it -- the step is then taken and the fit ends); a
// line search that finds no decrease in 16 backtracks ends the fit as stalled.
#include <math.h>
#include <string.h>
#include "cfit_dev.h"

extern "C" int qcs_net_forward_dev(qcs_net *nt, const qcs_params *p, const uint16_t *d_rois, int n, float *d_out, float *d_norm,
                                   void *stream);

namespace qcs {

// Work item j fits ROI roi = idx ? idx[j] : j (j < n, or < *n_dev when given).  Everything per
// ROI is indexed by roi: init (stride istride), theta_out (physical units, stride ostride),
// info (4 ints: iterations, flags, loss evaluations, 0) and ll (FP64 log-likelihood).
template <int K, int R>
__global__ void __launch_bounds__(64) cfit_kernel(const uint16_t *__restrict__ rois, const int *__restrict__ idx, int n,
                                                  const int *__restrict__ n_dev, const float *__restrict__ init, int istride,
                                                  float *__restrict__ theta_out, int ostride, int *__restrict__ info,
                                                  double *__restrict__ ll_out, CFitConst c)
{
    constexpr int P = CP<K>::P, NS = CP<K>::NS, RR = R * R, WR = R * (R + 1);
    __shared__ float qsh[RR * 64];
    __shared__ float Dsh[NS * 64];
    const int lane = threadIdx.x;
    const int j = blockIdx.x * 64 + lane;
    if (n_dev) n = *n_dev;
    if ((int)(blockIdx.x * 64) >= n) return;   // whole wave beyond the count
    const bool valid = j < n;
    const int roi = valid ? (idx ? idx[j] : j) : 0;
    float *qs = qsh + lane;
    float *D = Dsh + lane;
    {
        const uint16_t *src = rois + (size_t)roi * WR;
        for (int k = 0; k < RR; k++) {
            const float v = valid ? ((float)src[k] - c.offset) * c.kadc : 1.0f;
            qs[k * 64] = fmaxf(v, 1.0f) * c.inv_qe;
        }
    }
    float th[P], g[P];
    float zero[P + 1];
#pragma unroll
    for (int i = 0; i <= P; i++) zero[i] = 0.0f;
    cf_init<K, R>(qs, (valid && init) ? init + (size_t)roi * istride : zero, valid ? c.init_mode : CF_INIT_HEUR, c.sigma0, th);
    cf_project<K, R>(th);
    float L = cf_lossgrad<K, R, true>(qs, th, g);
    // starting inverse Hessian: the inverse Fisher information at the initial point (identity if
    // it is not positive definite; the first step is then normalised)
    bool first = !cf_fisher_inverse<K, R>(th, D);
    if (first) {
#pragma unroll
        for (int i = 0; i < P; i++)
#pragma unroll
            for (int jj = i; jj < P; jj++) D[tri(i, jj, P) * 64] = (i == jj) ? 1.0f : 0.0f;
    }
    int iters = 0, flags = 0, evals = 1;
    bool done = !valid;
    while (true) {
        if (__all(done)) break;   // wave-uniform
        // direction in scaled units: du = -D gu, gu = g * scale
        float gu[P], du[P];
#pragma unroll
        for (int i = 0; i < P; i++) gu[i] = g[i] * cf_scale(i, P);
        float slope = 0.0f;
#pragma unroll
        for (int i = 0; i < P; i++) {
            float a = 0.0f;
#pragma unroll
            for (int jj = 0; jj < P; jj++) a = fmaf(D[(i <= jj ? tri(i, jj, P) : tri(jj, i, P)) * 64], gu[jj], a);
            du[i] = -a;
            slope = fmaf(du[i], gu[i], slope);
        }
        if (!done) {
            float dmax = 0.0f;
#pragma unroll
            for (int i = 0; i < P; i++) dmax = fmaxf(dmax, fabsf(du[i]));
            if (dmax < c.tol && slope < 0.0f) {   // the full quasi-Newton step is below tolerance
                flags |= CF_CONVERGED;
                done = true;
            }
        }
        if (!(slope < 0.0f)) {   // not a descent direction: restart from the gradient
#pragma unroll
            for (int i = 0; i < P; i++) du[i] = -gu[i];
            slope = 0.0f;
#pragma unroll
            for (int i = 0; i < P; i++) slope = fmaf(du[i], gu[i], slope);
#pragma unroll
            for (int i = 0; i < P; i++)
#pragma unroll
                for (int jj = i; jj < P; jj++) D[tri(i, jj, P) * 64] = (i == jj) ? 1.0f : 0.0f;
            first = true;
        }
        if (first) {   // identity inverse Hessian: limit the first step to 0.1 scaled units
            float m = 0.0f;
#pragma unroll
            for (int i = 0; i < P; i++) m = fmaxf(m, fabsf(du[i]));
            if (m > 0.1f) {
                const float f = 0.1f / m;
                slope = 0.0f;
#pragma unroll
                for (int i = 0; i < P; i++) {
                    du[i] *= f;
                    slope = fmaf(du[i], gu[i], slope);
                }
            }
        }
        float dmax = 0.0f;
#pragma unroll
        for (int i = 0; i < P; i++) dmax = fmaxf(dmax, fabsf(du[i]));
        // Armijo backtracking (every active lane evaluates each trial together)
        float alpha = 1.0f, tn[P], Ln = L;
        bool floor_step = false;
#pragma unroll
        for (int i = 0; i < P; i++) tn[i] = th[i];
        bool accepted = done;
        for (int t = 0; t < 16; t++) {
            if (__all(accepted)) break;
            if (!accepted) {
#pragma unroll
                for (int i = 0; i < P; i++) tn[i] = fmaf(alpha, du[i] * cf_scale(i, P), th[i]);
                cf_project<K, R>(tn);
            }
            float gdum[P];
            const float Lt = cf_lossgrad<K, R, false>(qs, tn, gdum);
            if (!accepted) {
                evals++;
                if (Lt == Lt && Lt < L && Lt <= fmaf(1e-4f * alpha, slope, L)) {
                    accepted = true;
                    Ln = Lt;
                } else if (t == 0 && !(Lt < L) && dmax < 20.0f * c.tol) {
                    // the full step is tiny and the FP32 loss cannot resolve it: take it and stop
                    accepted = true;
                    floor_step = true;
                    Ln = Lt;
                } else {
                    alpha *= 0.3f;
                }
            }
        }
        if (!done && floor_step) {
            iters++;
#pragma unroll
            for (int i = 0; i < P; i++) th[i] = tn[i];
            flags |= CF_CONVERGED;
            done = true;
        }
        if (!done) {
            iters++;
            bool moved = accepted && Ln < L;
            float s[P];
#pragma unroll
            for (int i = 0; i < P; i++) s[i] = moved ? (tn[i] - th[i]) / cf_scale(i, P) : 0.0f;
            if (!moved) {
                flags |= CF_STALLED;
                done = true;
            } else {
                float gnew[P];
#pragma unroll
                for (int i = 0; i < P; i++) th[i] = tn[i];
                L = cf_lossgrad<K, R, true>(qs, th, gnew);
                evals++;
                float y[P], sy = 0.0f, yy = 0.0f;
#pragma unroll
                for (int i = 0; i < P; i++) {
                    y[i] = (gnew[i] - g[i]) * cf_scale(i, P);
                    sy = fmaf(s[i], y[i], sy);
                    yy = fmaf(y[i], y[i], yy);
                    g[i] = gnew[i];
                }
                if (sy > 1e-12f) {
                    if (first) {
                        const float sc = sy / yy;
#pragma unroll
                        for (int i = 0; i < P; i++)
#pragma unroll
                            for (int jj = i; jj < P; jj++) D[tri(i, jj, P) * 64] = (i == jj) ? sc : 0.0f;
                        first = false;
                    }
                    // D <- D - rho (s u^T + u s^T) + (rho^2 y.u + rho) s s^T, u = D y
                    const float rho = 1.0f / sy;
                    float u[P], yu = 0.0f;
#pragma unroll
                    for (int i = 0; i < P; i++) {
                        float a = 0.0f;
#pragma unroll
                        for (int jj = 0; jj < P; jj++) a = fmaf(D[(i <= jj ? tri(i, jj, P) : tri(jj, i, P)) * 64], y[jj], a);
                        u[i] = a;
                        yu = fmaf(y[i], a, yu);
                    }
                    const float cs = fmaf(rho * rho, yu, rho);
#pragma unroll
                    for (int i = 0; i < P; i++)
#pragma unroll
                        for (int jj = i; jj < P; jj++) {
                            float* e = &D[tri(i, jj, P) * 64];
                            *e = fmaf(cs * s[i], s[jj], fmaf(-rho, fmaf(s[i], u[jj], u[i] * s[jj]), *e));
                        }
                }
                if (iters >= c.max_iter) {
                    flags |= CF_MAXITER;
                    done = true;
                }
            }
        }
    }
    if (!valid) return;
#pragma unroll
    for (int i = 0; i < P; i++) theta_out[(size_t)roi * ostride + i] = th[i];
    if (info) {
        info[(size_t)roi * 4 + 0] = iters;
        info[(size_t)roi * 4 + 1] = flags;
        info[(size_t)roi * 4 + 2] = evals;
        info[(size_t)roi * 4 + 3] = 0;
    }
    if (ll_out) ll_out[roi] = cf_loglik64<K, R>(qs, th);
}

// Model selection (FP64 log-likelihood ratio) and rows: ROI j gets rows 3j .. 3j+2, the unused
// ones zero.  sel[j] receives the chosen emitter count (0: no valid model).
template <int R>
__global__ void __launch_bounds__(64) cf_select_kernel(const uint16_t *__restrict__ rois, const int *__restrict__ idx, int n,
                                                       const int *__restrict__ n_dev, const float *__restrict__ t1,
                                                       const float *__restrict__ t2,
                                                       const float *__restrict__ t3, const double *__restrict__ l1,
                                                       const double *__restrict__ l2, const double *__restrict__ l3,
                                                       CSelConst c, float *__restrict__ rows, int *__restrict__ sel)
{
    __shared__ double Ish[CP<3>::NS * 64];
    const int lane = threadIdx.x;
    const int j = blockIdx.x * 64 + lane;
    if (n_dev) n = *n_dev;
    if (j >= n) return;
    const int roi = idx ? idx[j] : j;
    const uint16_t *hdr = rois + (size_t)roi * R * (R + 1) + R * R;
    float *o = rows + (size_t)roi * 3 * QCS_REC_FLOATS;
    for (int i = 0; i < 3 * QCS_REC_FLOATS; i++) o[i] = 0.0f;
    int K = 0;
    if (cf_valid<1, R>(t1 + (size_t)roi * 5, c.n_min)) K = 1;
    if (c.kmax >= 2 && t2 && cf_valid<2, R>(t2 + (size_t)roi * 8, c.n_min) && (K == 0 || l2[roi] - l1[roi] > (double)c.llr_half))
        K = 2;
    if (c.kmax >= 3 && t3 && K == 2 && cf_valid<3, R>(t3 + (size_t)roi * 11, c.n_min) && l3[roi] - l2[roi] > (double)c.llr_half)
        K = 3;
    if (sel) sel[roi] = K;
    double *I = Ish + lane;
    if (K == 1) cf_write_rows<1, R>(t1 + (size_t)roi * 5, hdr, c, I, o);
    else if (K == 2) cf_write_rows<2, R>(t2 + (size_t)roi * 8, hdr, c, I, o);
    else if (K == 3) cf_write_rows<3, R>(t3 + (size_t)roi * 11, hdr, c, I, o);
}

CFitConst cfit_const(const qcs_params &p, const qcs_tuning &t, int K, int init_mode)
{
    CFitConst c;
    c.offset = p.Offset;
    c.kadc = p.KAdc;
    c.inv_qe = 1.0f / p.QE;
    c.max_iter = t.max_iter[(K >= 1 && K <= 3) ? K - 1 : 0];
    c.tol = t.tol;
    c.init_mode = init_mode;
    c.sigma0 = 1.3f;
    return c;
}

int cfit_launch(qcs_ctx *ctx, int K, const uint16_t *d_rois, const int *d_idx, int n, const int *d_n, int init_mode,
                const float *d_init, int istride, float *d_theta, int ostride, int *d_info, double *d_ll, hipStream_t s)
{
    if (n <= 0) return QCS_OK;
    const qcs_params &p = ctx->p;
    if (p.LocType != QCS_LOCTYPE_GS2D) return fail(QCS_ERR_ARG, "converged fitting supports the 2D model only");
    if (ctx->tune.queue) {
        CfqSet x;
        memset(&x, 0, sizeof(x));
        x.rois = d_rois;
        x.idx = d_idx;
        x.n_dev = d_n;
        x.init = d_init;
        x.theta = d_theta;
        x.info = d_info;
        x.ll = d_ll;
        x.n = n;
        x.R = p.ROISize;
        x.istride = istride;
        x.ostride = ostride;
        return cfq_launch(ctx, K, init_mode, &x, 1, s);
    }
    const CFitConst c = cfit_const(p, ctx->tune, K, init_mode);
    const dim3 grid((n + 63) / 64), block(64);
    const int R = p.ROISize;
    if (K == 1) {
        CF_DISPATCH_R(R, hipLaunchKernelGGL((cfit_kernel<1, RT>), grid, block, 0, s, d_rois, d_idx, n, d_n, d_init, istride, d_theta, ostride, d_info, d_ll, c))
    } else if (K == 2) {
        CF_DISPATCH_R(R, hipLaunchKernelGGL((cfit_kernel<2, RT>), grid, block, 0, s, d_rois, d_idx, n, d_n, d_init, istride, d_theta, ostride, d_info, d_ll, c))
    } else if (K == 3) {
        CF_DISPATCH_R(R, hipLaunchKernelGGL((cfit_kernel<3, RT>), grid, block, 0, s, d_rois, d_idx, n, d_n, d_init, istride, d_theta, ostride, d_info, d_ll, c))
    } else {
        return fail(QCS_ERR_ARG, "emitter count %d not in 1..3", K);
    }
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

// K = 1, 2, 3 in sequence over the listed ROIs (or all n); t1/t2/t3 (strides 5/8/11) and l1..l3
// are indexed by ROI
int plain_multi_launch(qcs_ctx *ctx, int kmax, const uint16_t *d_rois, const int *d_idx, int n, const int *d_n, float *t1,
                       float *t2, float *t3, double *l1, double *l2, double *l3, hipStream_t s)
{
    int rc;
    if ((rc = cfit_launch(ctx, 1, d_rois, d_idx, n, d_n, CF_INIT_HEUR, nullptr, 0, t1, 5, nullptr, l1, s))) return rc;
    if (kmax >= 2 && (rc = cfit_launch(ctx, 2, d_rois, d_idx, n, d_n, CF_INIT_ADD, t1, 5, t2, 8, nullptr, l2, s))) return rc;
    if (kmax >= 3 && (rc = cfit_launch(ctx, 3, d_rois, d_idx, n, d_n, CF_INIT_ADD, t2, 8, t3, 11, nullptr, l3, s))) return rc;
    return QCS_OK;
}

CSelConst csel_const(const qcs_params &p, int kmax)
{
    CSelConst c;
    c.llr_half = 8.14f;   // chi^2(3 dof) at p = 1e-3 is 16.27
    c.n_min = 50.0f;
    c.pixel_size = p.PixelSize;
    c.qe = p.QE;
    c.kmax = kmax;
    return c;
}

int cselect_launch(const qcs_params &p, int kmax, const uint16_t *d_rois, const int *d_idx, int n, const int *d_n, const float *t1,
                   const float *t2, const float *t3, const double *l1, const double *l2, const double *l3, float *d_rows,
                   int *d_sel, hipStream_t s)
{
    if (n <= 0) return QCS_OK;
    const CSelConst c = csel_const(p, kmax);
    const dim3 grid((n + 63) / 64), block(64);
    CF_DISPATCH_R(p.ROISize, hipLaunchKernelGGL((cf_select_kernel<RT>), grid, block, 0, s, d_rois, d_idx, n, d_n, t1, t2, t3, l1,
                                                 l2, l3, c, d_rows, d_sel))
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}


// ---- learned warm start -------------------------------------------------------------------
// Network outputs [n][16] (nn.hip): 3 count logits, (x/R, y/R, N/(2 pi scale)) for 3 slots,
// s - 1.3, (b - bhat)/scale.  The decode kernel turns them into a starting vector for the
// predicted emitter count and appends the ROI to that count's work list.
__global__ void warm_decode_kernel(const float *__restrict__ out, const float *__restrict__ norm, int n, int R, float pmin,
                                   float *__restrict__ init11, int *__restrict__ khat, int *__restrict__ cnt,
                                   int *__restrict__ lists, int *__restrict__ fb_list, int *__restrict__ info)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    int cls = -1;   // 0..2: warm list of count cls + 1; 3: fallback list
    if (j < n) {
        // the warm fits write info only for the ROIs they take
        *reinterpret_cast<int4 *>(info + (size_t)j * 4) = make_int4(0, 0, 0, 0);
        const float *o = out + (size_t)j * 16;
        int K = 1;
        if (o[1] > o[0]) K = 2;
        if (o[2] > o[K - 1]) K = 3;
        // probability of the predicted count; below pmin the ROI goes straight to the plain path
        const float m = o[K - 1];
        const float pk = 1.0f / (__expf(o[0] - m) + __expf(o[1] - m) + __expf(o[2] - m));
        if (pk >= pmin) {
            const float scale = norm[2 * j], bh = norm[2 * j + 1];
            float *t = init11 + (size_t)j * 11;
            for (int k = 0; k < K; k++) {
                t[3 * k] = fmaxf(o[5 + 3 * k] * 6.283185307179586f * scale, 30.0f);
                t[3 * k + 1] = fminf(fmaxf(o[3 + 3 * k] * (float)R, 0.3f), (float)R - 0.3f);
                t[3 * k + 2] = fminf(fmaxf(o[4 + 3 * k] * (float)R, 0.3f), (float)R - 0.3f);
            }
            t[3 * K] = fminf(fmaxf(o[12] + 1.3f, 0.6f), 3.0f);
            t[3 * K + 1] = fmaxf(bh + o[13] * scale, 0.5f);
            khat[j] = K;
            cls = K - 1;
        } else {
            khat[j] = 0;
            cls = 3;
        }
    }
    for (int c = 0; c < 4; c++) {
        const int pos = wave_append(&cnt[c], cls == c);
        if (pos >= 0) {
            if (c == 3) fb_list[pos] = j;
            else lists[(size_t)c * n + pos] = j;
        }
    }
}

// Fallback test of the warm fit of ROI j (count K = khat[j]): not converged, an invalid model, or
// a final loss above the loss at the plain heuristic start -> the ROI joins the fallback list.
template <int K, int R>
__device__ void warm_check_one(const float *qs, const float *t11, int flags, float n_min, float sigma0, bool &fallback)
{
    float th[CP<K>::P], h[CP<K>::P], g[CP<K>::P];
    for (int i = 0; i < CP<K>::P; i++) th[i] = t11[i];
    cf_init<K, R>(qs, nullptr, CF_INIT_HEUR, sigma0, h);
    cf_project<K, R>(h);
    const float Lw = cf_lossgrad<K, R, false>(qs, th, g);
    const float Lh = cf_lossgrad<K, R, false>(qs, h, g);
    fallback = !(flags & CF_CONVERGED) || !cf_valid<K, R>(t11, n_min) || !(Lw <= Lh);
}

template <int R>
__global__ void __launch_bounds__(64) warm_check_kernel(const uint16_t *__restrict__ rois, int n, const int *__restrict__ khat,
                                                        const float *__restrict__ t11, const int *__restrict__ info, CFitConst c,
                                                        float n_min, int *__restrict__ fb_cnt, int *__restrict__ fb_list)
{
    constexpr int RR = R * R, WR = R * (R + 1);
    __shared__ float qsh[RR * 64];
    const int lane = threadIdx.x, j = blockIdx.x * 64 + lane;
    if ((int)(blockIdx.x * 64) >= n) return;
    const bool valid = j < n;
    float *qs = qsh + lane;
    const uint16_t *src = rois + (size_t)(valid ? j : 0) * WR;
    for (int k = 0; k < RR; k++) qs[k * 64] = fmaxf(valid ? ((float)src[k] - c.offset) * c.kadc : 1.0f, 1.0f) * c.inv_qe;
    bool fb = false;
    const int K = valid ? khat[j] : 0;   // K = 0: already in the fallback list (low confidence)
    if (K > 0) {
        const int flags = info[(size_t)j * 4 + 1];
        fb = true;
        if (K == 1) warm_check_one<1, R>(qs, t11 + (size_t)j * 11, flags, n_min, c.sigma0, fb);
        else if (K == 2) warm_check_one<2, R>(qs, t11 + (size_t)j * 11, flags, n_min, c.sigma0, fb);
        else if (K == 3) warm_check_one<3, R>(qs, t11 + (size_t)j * 11, flags, n_min, c.sigma0, fb);
    }
    const int pos = wave_append(fb_cnt, fb);
    if (pos >= 0) fb_list[pos] = j;
}

// rows of the warm fits (3 per ROI, unused zero), FP64 full-Fisher CRLB
template <int R>
__global__ void __launch_bounds__(64) warm_rows_kernel(const uint16_t *__restrict__ rois, int n, const int *__restrict__ khat,
                                                       const float *__restrict__ t11, CSelConst c, float *__restrict__ rows,
                                                       int *__restrict__ sel)
{
    __shared__ double Ish[CP<3>::NS * 64];
    const int lane = threadIdx.x, j = blockIdx.x * 64 + lane;
    if (j >= n) return;
    const uint16_t *hdr = rois + (size_t)j * R * (R + 1) + R * R;
    float *o = rows + (size_t)j * 3 * QCS_REC_FLOATS;
    for (int i = 0; i < 3 * QCS_REC_FLOATS; i++) o[i] = 0.0f;
    const int K = khat[j];
    if (sel) sel[j] = K;
    double *I = Ish + lane;
    const float *t = t11 + (size_t)j * 11;
    if (K == 1) cf_write_rows<1, R>(t, hdr, c, I, o);
    else if (K == 2) cf_write_rows<2, R>(t, hdr, c, I, o);
    else if (K == 3) cf_write_rows<3, R>(t, hdr, c, I, o);
}

// training labels from the plain path: [K, (x, y, N) x 3 sorted by N descending, s, b, 0...]
__global__ void label_kernel(int n, const int *__restrict__ sel, const float *__restrict__ t1, const float *__restrict__ t2,
                             const float *__restrict__ t3, float *__restrict__ labels)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float *lb = labels + (size_t)j * 16;
    for (int i = 0; i < 16; i++) lb[i] = 0.0f;
    const int K = sel[j];
    if (K < 1 || K > 3) return;
    const float *t = K == 1 ? t1 + (size_t)j * 5 : (K == 2 ? t2 + (size_t)j * 8 : t3 + (size_t)j * 11);
    int ord[3] = {0, 1, 2};
    for (int a = 0; a < K; a++)
        for (int b = a + 1; b < K; b++)
            if (t[3 * ord[b]] > t[3 * ord[a]]) { const int x = ord[a]; ord[a] = ord[b]; ord[b] = x; }
    lb[0] = (float)K;
    for (int k = 0; k < K; k++) {
        lb[1 + 3 * k] = t[3 * ord[k] + 1];
        lb[2 + 3 * k] = t[3 * ord[k] + 2];
        lb[3 + 3 * k] = t[3 * ord[k]];
    }
    lb[10] = t[3 * K];
    lb[11] = t[3 * K + 1];
}

__global__ void warm_zero_kernel(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// per-ROI statistics of a path: [0..2] ROIs per count, [3] fallbacks, [4] iterations, [5] evaluations
__global__ void warm_stats_kernel(int n, const int *__restrict__ sel, const int *__restrict__ info, int stride_info,
                                  long long *__restrict__ st)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    long long it = 0, ev = 0;
    if (j < n && info) {
        it = info[(size_t)j * stride_info];
        ev = info[(size_t)j * stride_info + 2];
    }
    for (int off = 32; off > 0; off >>= 1) {
        it += __shfl_down(it, off, 64);
        ev += __shfl_down(ev, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd((unsigned long long *)&st[4], (unsigned long long)it);
        atomicAdd((unsigned long long *)&st[5], (unsigned long long)ev);
    }
    if (j < n && sel) {
        const int K = sel[j];
        if (K >= 1 && K <= 3) atomicAdd((unsigned long long *)&st[K - 1], 1ull);
    }
}

int warm_launch(qcs_ctx *ctx, qcs_net *net, const uint16_t *d_rois, int n, float *d_loc, int *d_sel, long long *h_stats,
                hipStream_t s)
{
    const qcs_params &p = ctx->p;
    const int R = p.ROISize;
    if (R != 7) return fail(QCS_ERR_ARG, "the warm start takes 7x7 ROIs");
    // scratch: out [n][16], norm [n][2], init/theta [n][11], khat, lists 3n, fb list n, counters,
    // info [n][4]; plain buffers for the fallback (roi-indexed)
    const size_t nf = (size_t)n * (16 + 2 + 11 + 5 + 8 + 11), ni = (size_t)n * (1 + 3 + 1 + 4) + 32 + 4;
    int rc;
    if ((rc = ctx->cf_theta.ensure(nf * 4)) || (rc = ctx->cf_aux.ensure(ni * 4 + 64)) || (rc = ctx->cf_ll.ensure((size_t)n * 3 * 8)) ||
        (rc = ctx->pin_cnt.ensure(128)))
        return rc;
    float *out = ctx->cf_theta.as<float>(), *norm = out + (size_t)n * 16, *t11 = norm + (size_t)n * 2;
    float *t1 = t11 + (size_t)n * 11, *t2 = t1 + (size_t)n * 5, *t3 = t2 + (size_t)n * 8;
    // counters first (8-byte aligned statistics): cnt[0..2] per predicted count, cnt[3] fallbacks
    int *cnt = ctx->cf_aux.as<int>();
    long long *st = reinterpret_cast<long long *>(cnt + 8);
    int *khat = cnt + 32, *lists = khat + n, *fbl = lists + (size_t)3 * n, *info = fbl + n;
    info = reinterpret_cast<int *>((reinterpret_cast<uintptr_t>(info) + 15) & ~(uintptr_t)15);   // 16-byte rows
    double *l1 = ctx->cf_ll.as<double>(), *l2 = l1 + n, *l3 = l2 + n;
    hipLaunchKernelGGL(warm_zero_kernel, dim3(1), dim3(64), 0, s, cnt, 8 + 12);
    hipEvent_t ev;
    if ((rc = timer_begin(ctx->timer, s, &ev))) return rc;
    if ((rc = qcs_net_forward_dev(net, &p, d_rois, n, out, norm, s))) return rc;
    if ((rc = timer_end(ctx->timer, TM_NN, s, ev, n))) return rc;
    if ((rc = timer_begin(ctx->timer, s, &ev))) return rc;
    hipLaunchKernelGGL(warm_decode_kernel, dim3((n + 255) / 256), dim3(256), 0, s, out, norm, n, R, ctx->tune.warm_pmin, t11, khat, cnt,
                       lists, fbl, info);
    QCS_HIP(hipGetLastError());
    for (int K = 1; K <= 3; K++)
        if ((rc = cfit_launch(ctx, K, d_rois, lists + (size_t)(K - 1) * n, n, cnt + K - 1, CF_INIT_GIVEN, t11, 11, t11, 11, info,
                              nullptr, s)))
            return rc;
    const CFitConst cc = cfit_const(p, ctx->tune, 1, CF_INIT_HEUR);
    const CSelConst sc = csel_const(p, 3);
    hipLaunchKernelGGL((warm_check_kernel<7>), dim3((n + 63) / 64), dim3(64), 0, s, d_rois, n, khat, t11, info, cc, sc.n_min,
                       cnt + 3, fbl);
    hipLaunchKernelGGL((warm_rows_kernel<7>), dim3((n + 63) / 64), dim3(64), 0, s, d_rois, n, khat, t11, sc, d_loc, d_sel);
    QCS_HIP(hipGetLastError());
    // fallback: the plain path over the listed ROIs (rows and counts overwrite the warm ones)
    if ((rc = plain_multi_launch(ctx, 3, d_rois, fbl, n, cnt + 3, t1, t2, t3, l1, l2, l3, s))) return rc;
    if ((rc = cselect_launch(p, 3, d_rois, fbl, n, cnt + 3, t1, t2, t3, l1, l2, l3, d_loc, d_sel, s))) return rc;
    if ((rc = timer_end(ctx->timer, TM_CFIT, s, ev, n))) return rc;
    if (h_stats) {
        hipLaunchKernelGGL(warm_stats_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, d_sel, info, 4, st);
        long long *h = ctx->pin_cnt.as<long long>();
        QCS_HIP(hipMemcpyAsync(h, st, 6 * sizeof(long long), hipMemcpyDeviceToHost, s));
        int *hc = reinterpret_cast<int *>(h + 8);
        QCS_HIP(hipMemcpyAsync(hc, cnt, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
        QCS_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < 6; i++) h_stats[i] = h[i];
        h_stats[3] = hc[3];
        h_stats[6] = hc[0];
        h_stats[7] = hc[1];
        h_stats[8] = hc[2];
    }
    return QCS_OK;
}

}  // namespace qcs

using namespace qcs;

extern "C" int qcs_cfit_dev(qcs_ctx *ctx, int K, const uint16_t *d_rois, int n, int init_mode, const float *d_init,
                            float *d_theta, int *d_info, double *d_ll, void *stream)
{
    if (!ctx || n < 0 || (n > 0 && (!d_rois || !d_theta)) || init_mode < 0 || init_mode > 2 ||
        (init_mode != CF_INIT_HEUR && !d_init))
        return fail(QCS_ERR_ARG, "invalid argument");
    if (init_mode == CF_INIT_ADD && K < 2) return fail(QCS_ERR_ARG, "adding an emitter needs K >= 2");
    hipEvent_t ev;
    int rc = timer_begin(ctx->timer, (hipStream_t)stream, &ev);
    if (rc) return rc;
    rc = cfit_launch(ctx, K, d_rois, nullptr, n, nullptr, init_mode, d_init, init_mode == CF_INIT_ADD ? 3 * K - 1 : 3 * K + 2, d_theta,
                     3 * K + 2, d_info, d_ll, (hipStream_t)stream);
    if (rc) return rc;
    return timer_end(ctx->timer, TM_CFIT, (hipStream_t)stream, ev, n);
}

extern "C" int qcs_fit_mle_dev(qcs_ctx *ctx, const uint16_t *d_rois, int n, int kmax, float *d_loc, int *d_sel, void *stream)
{
    if (!ctx || n < 0 || (n > 0 && (!d_rois || !d_loc)) || kmax < 1 || kmax > 3) return fail(QCS_ERR_ARG, "invalid argument");
    if (n == 0) return QCS_OK;
    hipStream_t s = (hipStream_t)stream;
    const size_t nf = (size_t)n * (5 + 8 + 11), nd = (size_t)n * 3;
    int rc;
    if ((rc = ctx->cf_theta.ensure(nf * sizeof(float))) || (rc = ctx->cf_ll.ensure(nd * sizeof(double)))) return rc;
    float *t1 = ctx->cf_theta.as<float>(), *t2 = t1 + (size_t)n * 5, *t3 = t2 + (size_t)n * 8;
    double *l1 = ctx->cf_ll.as<double>(), *l2 = l1 + n, *l3 = l2 + n;
    hipEvent_t ev;
    if ((rc = timer_begin(ctx->timer, s, &ev))) return rc;
    if ((rc = plain_multi_launch(ctx, kmax, d_rois, nullptr, n, nullptr, t1, t2, t3, l1, l2, l3, s))) return rc;
    if ((rc = cselect_launch(ctx->p, kmax, d_rois, nullptr, n, nullptr, t1, kmax >= 2 ? t2 : nullptr, kmax >= 3 ? t3 : nullptr, l1,
                             l2, l3, d_loc, d_sel, s)))
        return rc;
    return timer_end(ctx->timer, TM_CFIT, s, ev, n);
}

extern "C" int qcs_fit_mle_mixed_dev(qcs_ctx *ctx, const qcs_roi_set *sets, int nsets, int kmax, void *stream)
{
    if (!ctx || !sets || nsets < 1 || nsets > 3 || kmax < 1 || kmax > 3) return fail(QCS_ERR_ARG, "invalid argument");
    if (ctx->p.LocType != QCS_LOCTYPE_GS2D) return fail(QCS_ERR_ARG, "converged fitting supports the 2D model only");
    size_t total = 0;
    for (int i = 0; i < nsets; i++) {
        const qcs_roi_set &x = sets[i];
        if (x.roi_size != 7 && x.roi_size != 9 && x.roi_size != 11)
            return fail(QCS_ERR_ARG, "set %d: ROI size %d not in {7, 9, 11}", i, x.roi_size);
        if (x.n < 0 || (x.n > 0 && (!x.d_rois || !x.d_loc))) return fail(QCS_ERR_ARG, "set %d: invalid ROIs or rows", i);
        total += (size_t)x.n;
    }
    if (total == 0) return QCS_OK;
    if (total > (size_t)INT32_MAX / 32) return fail(QCS_ERR_ARG, "too many ROIs in one call");
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if ((rc = ctx->cf_theta.ensure(total * (5 + 8 + 11) * sizeof(float))) || (rc = ctx->cf_ll.ensure(total * 3 * sizeof(double))))
        return rc;
    // per set: t1 | t2 | t3 (strides 5, 8, 11) and l1 | l2 | l3, each indexed by ROI
    float *tb[3][3];
    double *lb[3][3];
    {
        float *t = ctx->cf_theta.as<float>();
        double *l = ctx->cf_ll.as<double>();
        for (int i = 0; i < nsets; i++) {
            const size_t m = (size_t)sets[i].n;
            tb[i][0] = t;
            tb[i][1] = t + m * 5;
            tb[i][2] = t + m * 13;
            t += m * 24;
            for (int k = 0; k < 3; k++) lb[i][k] = l + m * k;
            l += m * 3;
        }
    }
    hipEvent_t ev;
    if ((rc = timer_begin(ctx->timer, s, &ev))) return rc;
    for (int K = 1; K <= kmax; K++) {
        CfqSet xs[3];
        memset(xs, 0, sizeof(xs));
        for (int i = 0; i < nsets; i++) {
            CfqSet &x = xs[i];
            x.rois = sets[i].d_rois;
            x.n = sets[i].n;
            x.R = sets[i].roi_size;
            x.init = K == 1 ? nullptr : tb[i][K - 2];
            x.istride = K == 1 ? 0 : 3 * K - 1;
            x.theta = tb[i][K - 1];
            x.ostride = 3 * K + 2;
            x.ll = lb[i][K - 1];
        }
        if ((rc = cfq_launch(ctx, K, K == 1 ? CF_INIT_HEUR : CF_INIT_ADD, xs, nsets, s))) return rc;
    }
    for (int i = 0; i < nsets; i++) {
        qcs_params p = ctx->p;
        p.ROISize = sets[i].roi_size;
        if ((rc = cselect_launch(p, kmax, sets[i].d_rois, nullptr, sets[i].n, nullptr, tb[i][0], kmax >= 2 ? tb[i][1] : nullptr,
                                 kmax >= 3 ? tb[i][2] : nullptr, lb[i][0], lb[i][1], lb[i][2], sets[i].d_loc, sets[i].d_sel, s)))
            return rc;
    }
    return timer_end(ctx->timer, TM_CFIT, s, ev, (long long)total);
}

extern "C" int qcs_fit_warm_dev(qcs_ctx *ctx, qcs_net *net, const uint16_t *d_rois, int n, float *d_loc, int *d_sel,
                                long long *h_stats, void *stream)
{
    if (!ctx || !net || n < 0 || (n > 0 && (!d_rois || !d_loc))) return fail(QCS_ERR_ARG, "invalid argument");
    if (n == 0) return QCS_OK;
    if (!d_sel && h_stats) {   // the statistics count the selected emitters: keep them in scratch
        int rc = ctx->cf_sel.ensure((size_t)n * sizeof(int));
        if (rc) return rc;
        d_sel = ctx->cf_sel.as<int>();
    }
    return warm_launch(ctx, net, d_rois, n, d_loc, d_sel, h_stats, (hipStream_t)stream);
}

extern "C" int qcs_label_dev(qcs_ctx *ctx, const uint16_t *d_rois, int n, float *d_labels, int *d_sel, void *stream)
{
    if (!ctx || n < 0 || (n > 0 && (!d_rois || !d_labels))) return fail(QCS_ERR_ARG, "invalid argument");
    if (n == 0) return QCS_OK;
    hipStream_t s = (hipStream_t)stream;
    const size_t nf = (size_t)n * (5 + 8 + 11 + 36);
    int rc;
    if ((rc = ctx->cf_theta.ensure(nf * sizeof(float))) || (rc = ctx->cf_ll.ensure((size_t)n * 3 * sizeof(double))) ||
        (rc = ctx->cf_aux.ensure((size_t)n * sizeof(int) + 64)))
        return rc;
    float *t1 = ctx->cf_theta.as<float>(), *t2 = t1 + (size_t)n * 5, *t3 = t2 + (size_t)n * 8, *rows = t3 + (size_t)n * 11;
    double *l1 = ctx->cf_ll.as<double>(), *l2 = l1 + n, *l3 = l2 + n;
    int *sel = d_sel ? d_sel : ctx->cf_aux.as<int>();
    if ((rc = plain_multi_launch(ctx, 3, d_rois, nullptr, n, nullptr, t1, t2, t3, l1, l2, l3, s))) return rc;
    if ((rc = cselect_launch(ctx->p, 3, d_rois, nullptr, n, nullptr, t1, t2, t3, l1, l2, l3, rows, sel, s))) return rc;
    hipLaunchKernelGGL(label_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, sel, t1, t2, t3, d_labels);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

extern "C" int qcs_tuning_get(qcs_ctx *ctx, qcs_tuning *out)
{
    if (!ctx || !out) return fail(QCS_ERR_ARG, "null argument");
    *out = ctx->tune;
    return QCS_OK;
}

extern "C" int qcs_tuning_set(qcs_ctx *ctx, const qcs_tuning *t)
{
    if (!ctx || !t) return fail(QCS_ERR_ARG, "null argument");
    for (int k = 0; k < 3; k++)
        if (t->max_iter[k] < 1 || t->max_iter[k] > 1000) return fail(QCS_ERR_ARG, "max_iter must be in 1..1000");
    if (!(t->tol > 0.0f) || !(t->warm_pmin >= 0.0f && t->warm_pmin <= 1.0f)) return fail(QCS_ERR_ARG, "invalid tolerance or pmin");
    if ((t->queue != 0 && t->queue != 1) || t->waves_per_cu < 0 || t->waves_per_cu > 64)
        return fail(QCS_ERR_ARG, "queue must be 0 or 1 and waves_per_cu in 0..64");
    ctx->tune = *t;
    return QCS_OK;
}
