// cfit.hip -- converged multi-emitter Poisson-MLE fitting with FP64 model selection and FP64
// full-Fisher CRLB on gfx950.
//
// This is the "plain" path of the batched multi-emitter mode and the labeller of the learned
// warm start; it has no counterpart in the reference (whose fitters run a fixed iteration
// schedule, bfgsMLE_GS2D_core.h:313-426, bfgsMLE_AS3D_core.h).  Models, pixel (r, c) centred at
// (c + 0.5, r + 0.5), K = 1..3 emitters:
//
//   2D (LocType GS2D):    mu = sum_k N_k / (2 pi s^2) exp(-((c + .5 - x_k)^2 + (r + .5 - y_k)^2) / (2 s^2)) + b
//                         theta = [N_k, x_k, y_k]_k, s, b (3K + 2 parameters)
//   elliptical (AS3D):    mu = sum_k N_k / (2 pi sx_k sy_k) exp(-(dx^2 / (2 sx_k^2) + dy^2 / (2 sy_k^2))) + b
//                         theta = [N_k, x_k, y_k, sx_k, sy_k]_k, b (5K + 1 parameters), z from the
//                         calibration polynomial of sx^2 - sy^2 as in the reference's rows
//
// in photons and pixels; the optimiser works on u = theta / scale (N / 1000, b / 100) so the
// identity is a sensible first inverse Hessian.  Loss: Poisson negative log-likelihood sum(mu - q ln mu) with
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
template <int KK, int R>
__global__ void __launch_bounds__(64) cfit_kernel(const uint16_t *__restrict__ rois, const int *__restrict__ idx, int n,
                                                  const int *__restrict__ n_dev, const float *__restrict__ init, int istride,
                                                  float *__restrict__ theta_out, int ostride, int *__restrict__ info,
                                                  double *__restrict__ ll_out, CFitConst c)
{
    constexpr int P = CP<KK>::P, NS = CP<KK>::NS, RR = R * R, WR = R * (R + 1);
    __shared__ float qsh[RR * 64];
    __shared__ float Dsh[NS * 64];
    const int lane = threadIdx.x;
    const int j = blockIdx.x * 64 + lane;
    if (n_dev) n = *n_dev;
    if ((int)(blockIdx.x * 64) >= n) return;   // whole wave beyond the count
    const bool valid = j < n;
    const int roi = valid ? (idx ? idx[j] : j) : 0;
    float *qs = qsh + lane;
    DLds D{Dsh + lane};
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
    cf_init<KK, R>(qs, (valid && init) ? init + (size_t)roi * istride : zero, valid ? c.init_mode : CF_INIT_HEUR, c.sigma0, th);
    cf_project<KK, R>(th);
    float L = cf_lossgrad<KK, R, true>(qs, th, g);
    // starting inverse Hessian: the inverse Fisher information at the initial point (identity if
    // it is not positive definite; the first step is then normalised)
    bool first = !cf_fisher_inverse<KK, R>(th, D.p);
    if (first) {
#pragma unroll
        for (int i = 0; i < P; i++)
#pragma unroll
            for (int jj = i; jj < P; jj++) D[tri(i, jj, P)] = (i == jj) ? 1.0f : 0.0f;
    }
    int iters = 0, flags = 0, evals = 1, pin_run = 0;
    bool done = !valid;
    while (true) {
        if (__all(done)) break;   // wave-uniform
        float du[P], slope = 0.0f, dmax = 0.0f;
#pragma unroll
        for (int i = 0; i < P; i++) du[i] = 0.0f;
        if (!done && cf_direction<KK, R>(th, g, D, first, c.tol, c.active_set != 0, du, slope, dmax)) {
            flags |= CF_CONVERGED;
            done = true;
        }
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
                for (int i = 0; i < P; i++) tn[i] = fmaf(alpha, du[i] * cf_scale<KK>(i), th[i]);
                cf_project<KK, R>(tn);
            }
            float gdum[P];
            const float Lt = cf_lossgrad<KK, R, false>(qs, tn, gdum);
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
            const bool moved = accepted && Ln < L;
            if (!moved) {
                flags |= CF_STALLED;
                done = true;
            } else {
                float gnew[P];
                const float Lold = L;
                L = cf_lossgrad<KK, R, true>(qs, tn, gnew);
                evals++;
                cf_bfgs_update<KK>(D, first, th, tn, g, gnew);
#pragma unroll
                for (int i = 0; i < P; i++) {
                    th[i] = tn[i];
                    g[i] = gnew[i];
                }
                if constexpr (CM<KK>::AS && CM<KK>::K == 3)
                    if (c.pin_stop) pin_run = cf_at_bound<KK, R>(th) ? pin_run + 1 : 0;
                if (Lold - L < c.ftol) {   // the likelihood has stopped moving
                    flags |= CF_CONVERGED | CF_FLAT;
                    done = true;
                } else if (pin_run >= CF_PIN_RUN) {
                    flags |= CF_PINNED;
                    done = true;
                } else if (iters >= c.max_iter) {
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
    if (ll_out) ll_out[roi] = cf_loglik64<KK, R>(qs, th);
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
    c.ftol = K >= 2 ? t.ftol : 0.0f;   // one-emitter fits always run to the step tolerance
    c.active_set = t.active_set;
    c.pin_stop = (p.LocType == QCS_LOCTYPE_AS3D && K == 3) ? t.pin_stop : 0;
    if (p.LocType == QCS_LOCTYPE_AS3D) {
        // elliptical fits: bound-pinned parameters always leave the step, no flat-likelihood stop,
        // at least 100 iterations.  With five parameters per emitter the likelihood has long flat
        // valleys; on simulated two- and three-emitter ROIs (11 x 11) these settings raise the
        // two-emitter fits that end valid from 0.89 to 0.97 and the count accuracy at three
        // emitters from 0.63 to 0.81, and bring every valid fit within 1e-2 px of the FP64 optimum.
        c.ftol = 0.0f;
        c.active_set = 1;
        c.max_iter = c.max_iter > 100 ? c.max_iter : 100;
    }
    return c;
}

int cfit_launch(qcs_ctx *ctx, int K, const uint16_t *d_rois, const int *d_idx, int n, const int *d_n, int init_mode,
                const float *d_init, int istride, float *d_theta, int ostride, int *d_info, double *d_ll, hipStream_t s)
{
    if (n <= 0) return QCS_OK;
    const qcs_params &p = ctx->p;
    if (p.LocType != QCS_LOCTYPE_GS2D && p.LocType != QCS_LOCTYPE_AS3D)
        return fail(QCS_ERR_ARG, "converged fitting supports the 2D and astigmatic 3D models");
    if (K < 1 || K > 3) return fail(QCS_ERR_ARG, "emitter count %d not in 1..3", K);
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
        x.init_mode = init_mode;
        return cfq_launch(ctx, K, &x, 1, s);
    }
    const CFitConst c = cfit_const(p, ctx->tune, K, init_mode);
    const dim3 grid((n + 63) / 64), block(64);
    const int R = p.ROISize;
    CF_DISPATCH_KK(cf_kk(p, K), CF_DISPATCH_R(R, hipLaunchKernelGGL((cfit_kernel<KT, RT>), grid, block, 0, s, d_rois, d_idx, n, d_n, d_init,
                                                                  istride, d_theta, ostride, d_info, d_ll, c)))
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

// ---- staged model selection --------------------------------------------------------------------
// The emitter count follows the likelihood-ratio rule: K = 1 if the one-emitter fit is valid;
// K = 2 if the two-emitter fit is valid and (K = 0 or ll2 - ll1 > llr_half); K = 3 if K = 2, the
// three-emitter fit is valid and ll3 - ll2 > llr_half (valid: every emitter >= n_min photons
// inside the ROI, a plausible width).  K = 3 is reachable only from K = 2, so the three-emitter
// fit runs only for the ROIs the first stage leaves at K = 2; their list is built on the device.
CSelConst csel_const(const qcs_params &p, int kmax)
{
    CSelConst c;
    c.llr_half = 8.14f;   // chi^2(3 dof) at p = 1e-3 is 16.27
    c.n_min = 50.0f;
    c.pixel_size = p.PixelSize;
    c.qe = p.QE;
    c.kmax = kmax;
    c.pG[0] = p.p4_XGY; c.pG[1] = p.p3_XGY; c.pG[2] = p.p2_XGY; c.pG[3] = p.p1_XGY; c.pG[4] = p.p0_XGY;
    c.pL[0] = p.p4_XLY; c.pL[1] = p.p3_XLY; c.pL[2] = p.p2_XLY; c.pL[3] = p.p1_XLY; c.pL[4] = p.p0_XLY;
    c.zcorr = p.ZDepthCorrFactor;
    return c;
}

// stage 1: K in {0, 1, 2} into sel[roi]; kcap (may be null) caps the count per ROI (a cap of 1:
// the two-emitter fit was not run).  ROIs at K = 2 with kmax and cap >= 3 are appended to list3.
// E = 0 selects the 2D model, CF_ELLIPTIC the elliptical one (fit vectors t1, t2 of their sizes).
constexpr int SEL_WAVES = 16;   // work-group of the selection kernels (one list atomic per work-group)

template <int E, int R>
__global__ void __launch_bounds__(64 * SEL_WAVES) cf_select12_kernel(const int *__restrict__ idx, int n, const int *__restrict__ n_dev,
                                                         const float *__restrict__ t1, const float *__restrict__ t2,
                                                         const double *__restrict__ l1, const double *__restrict__ l2,
                                                         const int *__restrict__ kcap, CSelConst c, int *__restrict__ sel,
                                                         int *__restrict__ cnt3, int *__restrict__ list3)
{
    __shared__ BlockAppend<SEL_WAVES, 1> ba;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (n_dev) n = min(*n_dev, n);
    if ((int)(blockIdx.x * blockDim.x) >= n) return;   // whole work-group beyond the count
    int K = -1, roi = 0, cap = 3;
    if (j < n) {
        roi = idx ? idx[j] : j;
        if (kcap) cap = kcap[roi];
        K = 0;
        if (cf_valid<1 + E, R>(t1 + (size_t)roi * CM<1 + E>::P, c.n_min)) K = 1;
        const bool have2 = c.kmax >= 2 && t2 && cap >= 2;
        if (have2 && cf_valid<2 + E, R>(t2 + (size_t)roi * CM<2 + E>::P, c.n_min) && (K == 0 || l2[roi] - l1[roi] > (double)c.llr_half))
            K = 2;
        sel[roi] = K;
    }
    if (c.kmax >= 3) {
        const bool take[1] = {K == 2 && cap >= 3};
        int pos[1];
        block_append(ba, cnt3, take, pos);
        if (pos[0] >= 0) list3[pos[0]] = roi;
    }
}

// stage 2: K = 3 where it beats K = 2, final count into sel[roi], the unused rows of
// 3 roi .. 3 roi + 2 zeroed, and the ROI appended to the list of its count (klist + (K - 1) cap,
// counts in kcnt[0..2]): the rows kernels below each evaluate one model, so no wave runs the
// Fisher matrices of several emitter counts one after the other with lanes masked.
template <int E, int R>
__global__ void __launch_bounds__(64 * SEL_WAVES) cf_kpart_kernel(const int *__restrict__ idx, int n, const int *__restrict__ n_dev,
                                                      const float *__restrict__ t3, const double *__restrict__ l2,
                                                      const double *__restrict__ l3, CSelConst c, int *__restrict__ sel,
                                                      float *__restrict__ rows, int *__restrict__ kcnt, int *__restrict__ klist,
                                                      int cap)
{
    __shared__ BlockAppend<SEL_WAVES, 3> ba;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (n_dev) n = min(*n_dev, n);
    if ((int)(blockIdx.x * blockDim.x) >= n) return;   // whole work-group beyond the count
    int K = 0, roi = 0;
    if (j < n) {
        roi = idx ? idx[j] : j;
        K = sel[roi];
        if (K == 2 && c.kmax >= 3 && t3 && cf_valid<3 + E, R>(t3 + (size_t)roi * CM<3 + E>::P, c.n_min) &&
            l3[roi] - l2[roi] > (double)c.llr_half)
            K = 3;
        sel[roi] = K;
        float *o = rows + (size_t)roi * 3 * QCS_REC_FLOATS;
        if ((reinterpret_cast<uintptr_t>(rows) & 15) == 0) {   // 48-byte records: 16-byte stores
            for (int v = K * (QCS_REC_FLOATS / 4); v < 3 * (QCS_REC_FLOATS / 4); v++)
                reinterpret_cast<float4 *>(o)[v] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        } else {
            for (int i = K * QCS_REC_FLOATS; i < 3 * QCS_REC_FLOATS; i++) o[i] = 0.0f;
        }
    }
    const bool take[3] = {K == 1, K == 2, K == 3};
    int pos[3];
    block_append(ba, kcnt, take, pos);
#pragma unroll
    for (int k = 0; k < 3; k++)
        if (pos[k] >= 0) klist[(size_t)k * cap + pos[k]] = roi;
}

// rows 3 roi .. 3 roi + K - 1 of the listed ROIs of one model (KK) with the FP64 full-Fisher
// CRLB (the packed Fisher matrix in registers)
template <int KK, int R>
__global__ void __launch_bounds__(64) cf_rows_kernel(const uint16_t *__restrict__ rois, const int *__restrict__ list,
                                                     const int *__restrict__ cnt, int cap, const float *__restrict__ t,
                                                     CSelConst c, float *__restrict__ rows)
{
    const int j = blockIdx.x * 64 + threadIdx.x;
    if (j >= min(*cnt, cap)) return;
    const int roi = list[j];
    const uint16_t *hdr = rois + (size_t)roi * R * (R + 1) + R * R;
    cf_write_rows<KK, R>(t + (size_t)roi * CM<KK>::P, hdr, c, rows + (size_t)roi * 3 * QCS_REC_FLOATS);
}

int select12_launch(const qcs_params &p, int kmax, const int *idx, int n, const int *d_n, const float *t1, const float *t2,
                    const double *l1, const double *l2, const int *kcap, int *sel, int *cnt3, int *list3, hipStream_t s)
{
    if (n <= 0) return QCS_OK;
    const CSelConst c = csel_const(p, kmax);
    const dim3 grid((n + 64 * SEL_WAVES - 1) / (64 * SEL_WAVES)), block(64 * SEL_WAVES);
    if (p.LocType == QCS_LOCTYPE_AS3D) {
        CF_DISPATCH_R(p.ROISize, hipLaunchKernelGGL((cf_select12_kernel<CF_ELLIPTIC, RT>), grid, block, 0, s, idx, n, d_n, t1,
                                                     kmax >= 2 ? t2 : nullptr, l1, l2, kcap, c, sel, cnt3, list3))
    } else {
        CF_DISPATCH_R(p.ROISize, hipLaunchKernelGGL((cf_select12_kernel<0, RT>), grid, block, 0, s, idx, n, d_n, t1,
                                                     kmax >= 2 ? t2 : nullptr, l1, l2, kcap, c, sel, cnt3, list3))
    }
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

// final counts and rows of the n listed ROIs (all n when idx is null; *d_n caps n when given);
// kcnt: 3 device ints, klist: 3 n ints of scratch
int rows_launch(const qcs_params &p, int kmax, const uint16_t *rois, const int *idx, int n, const int *d_n, const float *t1,
                const float *t2, const float *t3, const double *l2, const double *l3, int *sel, float *rows, int *kcnt,
                int *klist, hipStream_t s)
{
    if (n <= 0) return QCS_OK;
    const CSelConst c = csel_const(p, kmax);
    const dim3 grid((n + 63) / 64), block(64);
    const dim3 sgrid((n + 64 * SEL_WAVES - 1) / (64 * SEL_WAVES)), sblock(64 * SEL_WAVES);
    fill_async(kcnt, 3, 0, s);
    const float *tk[3] = {t1, t2, t3};
    if (p.LocType == QCS_LOCTYPE_AS3D) {
        CF_DISPATCH_R(p.ROISize, hipLaunchKernelGGL((cf_kpart_kernel<CF_ELLIPTIC, RT>), sgrid, sblock, 0, s, idx, n, d_n,
                                                     kmax >= 3 ? t3 : nullptr, l2, l3, c, sel, rows, kcnt, klist, n))
        QCS_HIP(hipGetLastError());
        CF_DISPATCH_R(p.ROISize, {
            hipLaunchKernelGGL((cf_rows_kernel<1 + CF_ELLIPTIC, RT>), grid, block, 0, s, rois, klist, kcnt, n, tk[0], c, rows);
            if (kmax >= 2)
                hipLaunchKernelGGL((cf_rows_kernel<2 + CF_ELLIPTIC, RT>), grid, block, 0, s, rois, klist + n, kcnt + 1, n, tk[1],
                                   c, rows);
            if (kmax >= 3)
                hipLaunchKernelGGL((cf_rows_kernel<3 + CF_ELLIPTIC, RT>), grid, block, 0, s, rois, klist + 2 * (size_t)n,
                                   kcnt + 2, n, tk[2], c, rows);
        })
    } else {
        CF_DISPATCH_R(p.ROISize, hipLaunchKernelGGL((cf_kpart_kernel<0, RT>), sgrid, sblock, 0, s, idx, n, d_n,
                                                     kmax >= 3 ? t3 : nullptr, l2, l3, c, sel, rows, kcnt, klist, n))
        QCS_HIP(hipGetLastError());
        CF_DISPATCH_R(p.ROISize, {
            hipLaunchKernelGGL((cf_rows_kernel<1, RT>), grid, block, 0, s, rois, klist, kcnt, n, tk[0], c, rows);
            if (kmax >= 2)
                hipLaunchKernelGGL((cf_rows_kernel<2, RT>), grid, block, 0, s, rois, klist + n, kcnt + 1, n, tk[1], c, rows);
            if (kmax >= 3)
                hipLaunchKernelGGL((cf_rows_kernel<3, RT>), grid, block, 0, s, rois, klist + 2 * (size_t)n, kcnt + 2, n, tk[2],
                                   c, rows);
        })
    }
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

// The plain path over the listed ROIs (or all n): K = 1 from the heuristic start, K = 2 from the
// K = 1 fit plus an emitter at the largest residual, stage 1, K = 3 (likewise from K = 2) for the
// ROIs at K = 2, stage 2 and rows.  Fit vectors t1/t2/t3 (strides: the models' parameter counts,
// 5/8/11 in 2D, 6/11/16 elliptical), l1..l3, sel and the optional info arrays are indexed by ROI;
// cnt3 is one device int, list3 holds n ints.
int plain_path(qcs_ctx *ctx, int kmax, const uint16_t *rois, const int *idx, int n, const int *d_n, float *t1, float *t2,
               float *t3, double *l1, double *l2, double *l3, int *sel, int *cnt3, int *list3, float *rows, int *info1, int *info2,
               int *info3, int *kcnt, int *klist, hipStream_t s)
{
    int rc;
    const int n1 = cf_np(cf_kk(ctx->p, 1)), n2 = cf_np(cf_kk(ctx->p, 2)), n3 = cf_np(cf_kk(ctx->p, 3));
    if ((rc = cfit_launch(ctx, 1, rois, idx, n, d_n, CF_INIT_HEUR, nullptr, 0, t1, n1, info1, l1, s))) return rc;
    if (kmax >= 2 && (rc = cfit_launch(ctx, 2, rois, idx, n, d_n, CF_INIT_ADD, t1, n1, t2, n2, info2, l2, s))) return rc;
    fill_async(cnt3, 1, 0, s);
    if ((rc = select12_launch(ctx->p, kmax, idx, n, d_n, t1, t2, l1, l2, nullptr, sel, cnt3, list3, s))) return rc;
    if (kmax >= 3 && (rc = cfit_launch(ctx, 3, rois, list3, n, cnt3, CF_INIT_ADD, t2, n2, t3, n3, info3, l3, s))) return rc;
    return rows_launch(ctx->p, kmax, rois, idx, n, d_n, t1, t2, t3, l2, l3, sel, rows, kcnt, klist, s);
}

// ---- learned warm start -----------------------------------------------------------------------
// Network outputs [n][16] (nn.hip): 3 count logits, (x/R, y/R, N/(2 pi scale)) for 3 slots
// (brightest first), s - 1.3, (b - bhat)/scale.  The warm path keeps the selection rule of the
// plain path and changes where the fits start and which fits run:
//   - every ROI's one-emitter fit starts from the heuristic, as on the plain path;
//   - a count predicted with probability >= pmin caps the ROI's count: predicted single ROIs skip
//     the multi-emitter fits, predicted doubles the three-emitter fit;
//   - the others get the two-emitter fit, started from the network's two brightest slots where it
//     predicts two or more emitters and from the plain start otherwise;
//   - the three-emitter fit runs, from the plain start, for the uncapped ROIs stage 1 leaves at
//     K = 2.
// A warm fit that does not converge, or ends at a higher loss than its plain start has, is redone
// from the plain start (the fallback).
enum { WC_L2W = 0, WC_L2A = 1, WC_F1 = 2, WC_F2 = 3, WC_L3 = 4, WC_N = 16 };

__global__ void __launch_bounds__(64 * SEL_WAVES)
    warm_decode_kernel(const float *__restrict__ out, const float *__restrict__ norm, int n, int R, float pmin, float *__restrict__ tw1,
                       float *__restrict__ tw2, int *__restrict__ cs, int *__restrict__ khat, int *__restrict__ cnt,
                       int *__restrict__ l2w, int *__restrict__ l2a)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    int cls = -1;   // 0: two-emitter fit from the network start; 1: from the plain start
    if (j < n) {
        const float *o = out + (size_t)j * 16;
        int K = 1;
        if (o[1] > o[0]) K = 2;
        if (o[2] > o[K - 1]) K = 3;
        const float m = o[K - 1];
        const float pk = 1.0f / (__expf(o[0] - m) + __expf(o[1] - m) + __expf(o[2] - m));
        const float scale = norm[2 * j], bh = norm[2 * j + 1];
        const float s0 = fminf(fmaxf(o[12] + 1.3f, 0.6f), 3.0f), b0 = fmaxf(bh + o[13] * scale, 0.5f);
        float slot[2][3];
        for (int k = 0; k < 2; k++) {
            slot[k][0] = fmaxf(o[5 + 3 * k] * 6.283185307179586f * scale, 30.0f);
            slot[k][1] = fminf(fmaxf(o[3 + 3 * k] * (float)R, 0.3f), (float)R - 0.3f);
            slot[k][2] = fminf(fmaxf(o[4 + 3 * k] * (float)R, 0.3f), (float)R - 0.3f);
        }
        float *t = tw1 + (size_t)j * 5;
        t[0] = slot[0][0];
        t[1] = slot[0][1];
        t[2] = slot[0][2];
        t[3] = s0;
        t[4] = b0;
        khat[j] = K;
        const bool single = K == 1 && pk >= pmin;
        cs[j] = pk >= pmin ? K : 3;   // per-ROI count cap
        if (!single) {
            if (K >= 2) {
                float *u = tw2 + (size_t)j * 8;
                for (int k = 0; k < 2; k++)
                    for (int e = 0; e < 3; e++) u[3 * k + e] = slot[k][e];
                u[6] = s0;
                u[7] = b0;
                cls = 0;
            } else {
                cls = 1;
            }
        }
    }
    __shared__ BlockAppend<SEL_WAVES, 2> ba;
    static_assert(WC_L2A == WC_L2W + 1, "the two start lists' counters are adjacent");
    const bool take[2] = {cls == 0, cls == 1};
    int pos[2];
    block_append(ba, cnt + WC_L2W, take, pos);
    if (pos[0] >= 0) l2w[pos[0]] = j;
    if (pos[1] >= 0) l2a[pos[1]] = j;
}

// fallback test of the warm K-emitter fits of the listed ROIs (or all n): not converged, or a loss
// above the loss at the plain start (K = 1: heuristic; K = 2: the K = 1 fit t_prev plus an emitter
// at the largest residual) -> the ROI is appended to flist
template <int K, int R>
__global__ void __launch_bounds__(64) warm_check_kernel(const uint16_t *__restrict__ rois, const int *__restrict__ idx, int n,
                                                        const int *__restrict__ n_dev, const float *__restrict__ t_fit,
                                                        const int *__restrict__ info, const float *__restrict__ t_prev, CFitConst c,
                                                        int *__restrict__ fcnt, int *__restrict__ flist)
{
    constexpr int P = CP<K>::P;
    __shared__ float qsh[R * R * 64];
    const int lane = threadIdx.x, j = blockIdx.x * 64 + lane;
    if (n_dev) n = min(*n_dev, n);
    if ((int)(blockIdx.x * 64) >= n) return;   // whole wave beyond the count
    bool fb = false;
    int roi = 0;
    if (j < n) {
        roi = idx ? idx[j] : j;
        float *qs = qsh + lane;
        cf_load_q<R>(rois, roi, c, qs);
        float th[P], h[P], g[P];
        for (int i = 0; i < P; i++) th[i] = t_fit[(size_t)roi * P + i];
        cf_init<K, R>(qs, K == 1 ? nullptr : t_prev + (size_t)roi * 5, K == 1 ? CF_INIT_HEUR : CF_INIT_ADD, c.sigma0, h);
        cf_project<K, R>(h);
        const float Lw = cf_lossgrad<K, R, false>(qs, th, g);
        const float Lh = cf_lossgrad<K, R, false>(qs, h, g);
        fb = !(info[(size_t)roi * 4 + 1] & CF_CONVERGED) || !(Lw <= Lh);
    }
    const int pos = wave_append(fcnt, fb);
    if (pos >= 0) flist[pos] = roi;
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

// statistics of a multi-emitter call: [0..2] ROIs per final count, [4] iterations and [5] loss
// evaluations over the fits (info1..info3), [6] ROIs whose count the network capped below 3,
// [7] and [8] ROIs the network put at two and three emitters
__global__ void __launch_bounds__(64 * SEL_WAVES) cf_stats_kernel(int n, const int *__restrict__ sel, const int *__restrict__ info1, const int *__restrict__ info2,
                                const int *__restrict__ info3, const int *__restrict__ cs, const int *__restrict__ khat,
                                long long *__restrict__ st)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    long long v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (j < n) {
        const int K = sel[j];
        if (K >= 1 && K <= 3) v[K - 1] = 1;
        const int *inf[3] = {info1, info2, info3};
        for (int k = 0; k < 3; k++)
            if (inf[k]) {
                v[4] += inf[k][(size_t)j * 4];
                v[5] += inf[k][(size_t)j * 4 + 2];
            }
        if (cs) v[6] = cs[j] < 3;
        if (khat) {
            v[7] = khat[j] == 2;
            v[8] = khat[j] == 3;
        }
    }
    // wave sums, then one atomic per statistic and work-group
    __shared__ long long ws[SEL_WAVES][9];
    for (int i = 0; i < 9; i++) {
        long long x = v[i];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6][i] = x;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        long long x = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) x += ws[w][threadIdx.x];
        if (x) atomicAdd((unsigned long long *)&st[threadIdx.x], (unsigned long long)x);
    }
}

// ROI-indexed scratch of the multi-emitter paths, carved from the context's buffers
struct MeScratch {
    float *t1, *t2, *t3;
    double *l1, *l2, *l3;
    int *cnt;                  // WC_N counters
    int *sel, *list3;
    int *kcnt, *klist;         // rows: 3 counters per ROI set (cnt + 16 + 4 i), 3 n ints of lists
    int *info1, *info2, *info3;
    // warm path only
    float *out, *norm, *tw1, *tw2;
    int *cs, *khat, *l2w, *l2a, *f1, *f2;
    long long *st;             // 9 statistics (8-byte aligned)
};

int me_scratch(qcs_ctx *ctx, int n, bool warm, bool info, MeScratch &m)
{
    const size_t n4 = ((size_t)n + 3) & ~(size_t)3;   // keeps every int array 16-byte aligned
    // fit vectors sized for the larger (elliptical) model: 6 + 11 + 16 floats per ROI
    const size_t nf = n4 * (6 + 11 + 16) + (warm ? n4 * (16 + 2 + 5 + 8) : 0);
    const size_t ni = 64 + n4 * 5 + (info ? n4 * 12 : 0) + (warm ? n4 * 6 : 0);
    int rc;
    if ((rc = ctx->cf_theta.ensure(nf * sizeof(float))) || (rc = ctx->cf_ll.ensure(n4 * 3 * sizeof(double))) ||
        (rc = ctx->cf_list.ensure(ni * sizeof(int))))
        return rc;
    memset(&m, 0, sizeof(m));
    float *f = ctx->cf_theta.as<float>();
    m.t1 = f;
    m.t2 = f + n4 * 6;
    m.t3 = f + n4 * 17;
    f += n4 * 33;
    double *d = ctx->cf_ll.as<double>();
    m.l1 = d;
    m.l2 = d + n4;
    m.l3 = d + 2 * n4;
    int *p = ctx->cf_list.as<int>();
    m.cnt = p;
    m.kcnt = p + 16;
    m.st = reinterpret_cast<long long *>(p + 32);
    p += 64;
    m.sel = p;
    m.list3 = p + n4;
    m.klist = p + 2 * n4;
    p += 5 * n4;
    if (info) {
        m.info1 = p;
        m.info2 = p + 4 * n4;
        m.info3 = p + 8 * n4;
        p += 12 * n4;
    }
    if (warm) {
        m.out = f;
        m.norm = f + n4 * 16;
        m.tw1 = f + n4 * 18;
        m.tw2 = f + n4 * 23;
        m.cs = p;
        m.khat = p + n4;
        m.l2w = p + 2 * n4;
        m.l2a = p + 3 * n4;
        m.f1 = p + 4 * n4;
        m.f2 = p + 5 * n4;
    }
    return QCS_OK;
}

int stats_copy(qcs_ctx *ctx, const MeScratch &m, int n, const int *sel, long long *h_stats, hipStream_t s)
{
    int rc;
    if ((rc = ctx->pin_cnt.ensure(9 * sizeof(long long) + WC_N * sizeof(int)))) return rc;
    hipLaunchKernelGGL(cf_stats_kernel, dim3((n + 64 * SEL_WAVES - 1) / (64 * SEL_WAVES)), dim3(64 * SEL_WAVES), 0, s, n, sel, m.info1, m.info2, m.info3, m.cs, m.khat, m.st);
    QCS_HIP(hipGetLastError());
    long long *h = ctx->pin_cnt.as<long long>();
    int *hc = reinterpret_cast<int *>(h + 9);
    QCS_HIP(hipMemcpyAsync(h, m.st, 9 * sizeof(long long), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipMemcpyAsync(hc, m.cnt, WC_N * sizeof(int), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < 9; i++) h_stats[i] = h[i];
    h_stats[3] = (long long)hc[WC_F1] + hc[WC_F2];
    return QCS_OK;
}

int warm_launch(qcs_ctx *ctx, qcs_net *net, const uint16_t *d_rois, int n, float *d_loc, int *d_sel, long long *h_stats,
                hipStream_t s)
{
    const qcs_params &p = ctx->p;
    const int R = p.ROISize;
    if (R != 7) return fail(QCS_ERR_ARG, "the warm start takes 7x7 ROIs");
    if (p.LocType != QCS_LOCTYPE_GS2D) return fail(QCS_ERR_ARG, "the warm start supports the 2D model only");
    MeScratch m;
    int rc;
    if ((rc = me_scratch(ctx, n, true, true, m))) return rc;   // the fallback tests read the fits' flags
    int *sel = d_sel ? d_sel : m.sel;
    int *cnt = m.cnt;
    fill_async(cnt, 64, 0, s);   // counters and statistics
    if (h_stats) fill_async(m.info1, (size_t)12 * (((size_t)n + 3) & ~(size_t)3), 0, s);
    hipEvent_t ev;
    if ((rc = timer_begin(ctx->timer, s, &ev))) return rc;
    if ((rc = qcs_net_forward_dev(net, &p, d_rois, n, m.out, m.norm, s))) return rc;
    if ((rc = timer_end(ctx->timer, TM_NN, s, ev, n))) return rc;
    if ((rc = timer_begin(ctx->timer, s, &ev))) return rc;
    hipLaunchKernelGGL(warm_decode_kernel, dim3((n + 64 * SEL_WAVES - 1) / (64 * SEL_WAVES)), dim3(64 * SEL_WAVES), 0, s, m.out, m.norm, n, R, ctx->tune.warm_pmin, m.tw1,
                       m.tw2, m.cs, m.khat, cnt, m.l2w, m.l2a);
    QCS_HIP(hipGetLastError());
    const CFitConst c2 = cfit_const(p, ctx->tune, 2, CF_INIT_ADD);
    const dim3 grid((n + 63) / 64), block(64);
    // one-emitter fits from the heuristic start, as on the plain path: their time is the tail of
    // wrong-model fits of multi-emitter ROIs, which a warm start does not shorten, and a plain
    // start needs no fallback pass
    if ((rc = cfit_launch(ctx, 1, d_rois, nullptr, n, nullptr, CF_INIT_HEUR, nullptr, 0, m.t1, 5, m.info1, m.l1, s))) return rc;
    // two-emitter fits of the ROIs not confidently single
    if (ctx->tune.queue) {   // both starts in one persistent launch
        CfqSet xs[2];
        memset(xs, 0, sizeof(xs));
        for (int i = 0; i < 2; i++) {
            xs[i].rois = d_rois;
            xs[i].idx = i == 0 ? m.l2w : m.l2a;
            xs[i].n_dev = cnt + (i == 0 ? WC_L2W : WC_L2A);
            xs[i].init = i == 0 ? m.tw2 : m.t1;
            xs[i].istride = i == 0 ? 8 : 5;
            xs[i].init_mode = i == 0 ? CF_INIT_GIVEN : CF_INIT_ADD;
            xs[i].theta = m.t2;
            xs[i].ostride = 8;
            xs[i].info = m.info2;
            xs[i].ll = m.l2;
            xs[i].n = n;
            xs[i].R = R;
        }
        if ((rc = cfq_launch(ctx, 2, xs, 2, s))) return rc;
    } else {
        if ((rc = cfit_launch(ctx, 2, d_rois, m.l2w, n, cnt + WC_L2W, CF_INIT_GIVEN, m.tw2, 8, m.t2, 8, m.info2, m.l2, s))) return rc;
        if ((rc = cfit_launch(ctx, 2, d_rois, m.l2a, n, cnt + WC_L2A, CF_INIT_ADD, m.t1, 5, m.t2, 8, m.info2, m.l2, s))) return rc;
    }
    hipLaunchKernelGGL((warm_check_kernel<2, 7>), grid, block, 0, s, d_rois, m.l2w, n, cnt + WC_L2W, m.t2, m.info2, m.t1, c2,
                       cnt + WC_F2, m.f2);
    QCS_HIP(hipGetLastError());
    if ((rc = cfit_launch(ctx, 2, d_rois, m.f2, n, cnt + WC_F2, CF_INIT_ADD, m.t1, 5, m.t2, 8, m.info2, m.l2, s))) return rc;
    // selection as on the plain path
    if ((rc = select12_launch(p, 3, nullptr, n, nullptr, m.t1, m.t2, m.l1, m.l2, m.cs, sel, cnt + WC_L3, m.list3, s))) return rc;
    if ((rc = cfit_launch(ctx, 3, d_rois, m.list3, n, cnt + WC_L3, CF_INIT_ADD, m.t2, 8, m.t3, 11, m.info3, m.l3, s))) return rc;
    if ((rc = rows_launch(p, 3, d_rois, nullptr, n, nullptr, m.t1, m.t2, m.t3, m.l2, m.l3, sel, d_loc, m.kcnt, m.klist, s)))
        return rc;
    if ((rc = timer_end(ctx->timer, TM_CFIT, s, ev, n))) return rc;
    if (h_stats) return stats_copy(ctx, m, n, sel, h_stats, s);
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
    if (K < 1 || K > 3) return fail(QCS_ERR_ARG, "emitter count %d not in 1..3", K);
    hipEvent_t ev;
    int rc = timer_begin(ctx->timer, (hipStream_t)stream, &ev);
    if (rc) return rc;
    const int np = cf_np(cf_kk(ctx->p, K)), np0 = K > 1 ? cf_np(cf_kk(ctx->p, K - 1)) : 0;
    rc = cfit_launch(ctx, K, d_rois, nullptr, n, nullptr, init_mode, d_init, init_mode == CF_INIT_ADD ? np0 : np, d_theta, np, d_info,
                     d_ll, (hipStream_t)stream);
    if (rc) return rc;
    return timer_end(ctx->timer, TM_CFIT, (hipStream_t)stream, ev, n);
}

extern "C" int qcs_fit_mle_dev(qcs_ctx *ctx, const uint16_t *d_rois, int n, int kmax, float *d_loc, int *d_sel, void *stream)
{
    if (!ctx || n < 0 || (n > 0 && (!d_rois || !d_loc)) || kmax < 1 || kmax > 3) return fail(QCS_ERR_ARG, "invalid argument");
    if (n == 0) return QCS_OK;
    hipStream_t s = (hipStream_t)stream;
    MeScratch m;
    int rc;
    if ((rc = me_scratch(ctx, n, false, false, m))) return rc;
    hipEvent_t ev;
    if ((rc = timer_begin(ctx->timer, s, &ev))) return rc;
    if ((rc = plain_path(ctx, kmax, d_rois, nullptr, n, nullptr, m.t1, m.t2, m.t3, m.l1, m.l2, m.l3, d_sel ? d_sel : m.sel, m.cnt,
                         m.list3, d_loc, nullptr, nullptr, nullptr, m.kcnt, m.klist, s)))
        return rc;
    return timer_end(ctx->timer, TM_CFIT, s, ev, n);
}

extern "C" int qcs_fit_mle_mixed_dev(qcs_ctx *ctx, const qcs_roi_set *sets, int nsets, int kmax, void *stream)
{
    if (!ctx || !sets || nsets < 1 || nsets > 3 || kmax < 1 || kmax > 3) return fail(QCS_ERR_ARG, "invalid argument");
    if (ctx->p.LocType != QCS_LOCTYPE_GS2D && ctx->p.LocType != QCS_LOCTYPE_AS3D)
        return fail(QCS_ERR_ARG, "converged fitting supports the 2D and astigmatic 3D models");
    const int n1 = cf_np(cf_kk(ctx->p, 1)), n2 = cf_np(cf_kk(ctx->p, 2)), n3 = cf_np(cf_kk(ctx->p, 3));
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
    // one scratch carve for all sets; set i uses the ROI range [off_i, off_i + n_i)
    MeScratch m;
    int rc;
    if ((rc = me_scratch(ctx, (int)total, false, false, m))) return rc;
    size_t off[3] = {0, 0, 0};
    for (int i = 1; i < nsets; i++) off[i] = off[i - 1] + (size_t)sets[i - 1].n;
    hipEvent_t ev;
    if ((rc = timer_begin(ctx->timer, s, &ev))) return rc;
    fill_async(m.cnt, 8, 0, s);
    for (int K = 1; K <= kmax; K++) {
        CfqSet xs[3];
        memset(xs, 0, sizeof(xs));
        for (int i = 0; i < nsets; i++) {
            CfqSet &x = xs[i];
            const size_t o = off[i];
            x.rois = sets[i].d_rois;
            x.n = sets[i].n;
            x.n_dev = sets[i].d_n;
            x.R = sets[i].roi_size;
            x.init = K == 1 ? nullptr : (K == 2 ? m.t1 + o * n1 : m.t2 + o * n2);
            x.istride = K == 1 ? 0 : (K == 2 ? n1 : n2);
            x.theta = K == 1 ? m.t1 + o * n1 : (K == 2 ? m.t2 + o * n2 : m.t3 + o * n3);
            x.ostride = K == 1 ? n1 : (K == 2 ? n2 : n3);
            x.ll = (K == 1 ? m.l1 : (K == 2 ? m.l2 : m.l3)) + o;
            x.init_mode = K == 1 ? CF_INIT_HEUR : CF_INIT_ADD;
            if (K == 3) {   // only the ROIs that stage 1 left at two emitters
                x.idx = m.list3 + o;
                x.n_dev = m.cnt + i;
            }
        }
        if ((rc = cfq_launch(ctx, K, xs, nsets, s))) return rc;
        if (K == 2 || (K == 1 && kmax == 1)) {
            for (int i = 0; i < nsets; i++) {
                qcs_params p = ctx->p;
                p.ROISize = sets[i].roi_size;
                const size_t o = off[i];
                int *sel = sets[i].d_sel ? sets[i].d_sel : m.sel + o;
                if ((rc = select12_launch(p, kmax, nullptr, sets[i].n, sets[i].d_n, m.t1 + o * n1, m.t2 + o * n2, m.l1 + o,
                                          m.l2 + o, nullptr, sel, m.cnt + i, m.list3 + o, s)))
                    return rc;
            }
        }
    }
    for (int i = 0; i < nsets; i++) {
        qcs_params p = ctx->p;
        p.ROISize = sets[i].roi_size;
        const size_t o = off[i];
        int *sel = sets[i].d_sel ? sets[i].d_sel : m.sel + o;
        if ((rc = rows_launch(p, kmax, sets[i].d_rois, nullptr, sets[i].n, sets[i].d_n, m.t1 + o * n1, m.t2 + o * n2,
                              m.t3 + o * n3, m.l2 + o, m.l3 + o, sel, sets[i].d_loc, m.kcnt + 4 * i,
                              m.klist + 3 * o, s)))
            return rc;
    }
    return timer_end(ctx->timer, TM_CFIT, s, ev, (long long)total);
}

namespace qcs {

int extract_mixed_launch(qcs_ctx *ctx, const uint16_t *d_frames, int nframes, int start_frame, const int *d_start, int cap,
                         const qcs_roi_set *sets, int *cand_lists[3], int **cand_info, hipStream_t s);

// rows kept per candidate: cnt[cand[j]] = sel[j] for the set's ROIs j < *n
__global__ void __launch_bounds__(256) frame_rows_count_kernel(const int *__restrict__ cand, const int *__restrict__ n_dev, int n,
                                                               const int *__restrict__ sel, int *__restrict__ cnt)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || j >= *n_dev) return;
    cnt[cand[j]] = sel[j];
}

// the chosen emitters' rows of ROI j (rows 3j .. 3j + sel - 1) at out[off[cand[j]] ...]
__global__ void __launch_bounds__(256) frame_rows_gather_kernel(const int *__restrict__ cand, const int *__restrict__ n_dev, int n,
                                                                const int *__restrict__ sel, const int *__restrict__ off,
                                                                const float *__restrict__ rows, int out_cap, float *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || j >= *n_dev) return;
    const int k = sel[j], o = off[cand[j]];
    for (int e = 0; e < k; e++) {
        if (o + e >= out_cap) break;
        const float4 *src = reinterpret_cast<const float4 *>(rows + ((size_t)3 * j + e) * QCS_REC_FLOATS);
        float4 *dst = reinterpret_cast<float4 *>(out + (size_t)(o + e) * QCS_REC_FLOATS);
        dst[0] = src[0];
        dst[1] = src[1];
        dst[2] = src[2];
    }
}

// [0] rows written, [1] rows produced, [2] ROIs fitted (the three sets), [3] candidates found
__global__ void frame_rows_total_kernel(const int *__restrict__ tile_sums, int ntiles, int out_cap, const int *__restrict__ set_n,
                                        const int *__restrict__ info, int *__restrict__ n_out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int t = tile_sums[ntiles];
    n_out[0] = t < out_cap ? t : out_cap;
    n_out[1] = t;
    n_out[2] = set_n[0] + set_n[1] + set_n[2];
    n_out[3] = info[1];
}

// qcs_loc_mle_frames_dev; with d_start set the first frame number is read from it at run time
int loc_mle_frames(qcs_ctx *ctx, const uint16_t *d_frames, int nframes, int start_frame, const int *d_start, int cap, int kmax,
                   float *d_out, int out_cap, int *d_n_out, hipStream_t s);

}  // namespace qcs

extern "C" int qcs_loc_mle_frames_dev(qcs_ctx *ctx, const uint16_t *d_frames, int nframes, int start_frame, int cap, int kmax,
                                      float *d_out, int out_cap, int *d_n_out, void *stream)
{
    return loc_mle_frames(ctx, d_frames, nframes, start_frame, nullptr, cap, kmax, d_out, out_cap, d_n_out, (hipStream_t)stream);
}

namespace qcs {

int loc_mle_frames(qcs_ctx *ctx, const uint16_t *d_frames, int nframes, int start_frame, const int *d_start, int cap, int kmax,
                   float *d_out, int out_cap, int *d_n_out, hipStream_t s)
{
    if (!ctx || nframes < 0 || cap < 0 || (nframes > 0 && !d_frames) || kmax < 1 || kmax > 3 || out_cap < 0 ||
        (out_cap > 0 && !d_out) || !d_n_out)
        return fail(QCS_ERR_ARG, "invalid argument");
    if (ctx->p.LocType != QCS_LOCTYPE_GS2D && ctx->p.LocType != QCS_LOCTYPE_AS3D)
        return fail(QCS_ERR_ARG, "converged fitting supports the 2D and astigmatic 3D models");
    if ((long long)cap * 3 > (long long)INT32_MAX / 32) return fail(QCS_ERR_ARG, "capacity too large");
    const int ncap = cap > 0 ? cap : 1;
    if (nframes == 0 || cap == 0) {
        fill_async(d_n_out, 4, 0, s);
        QCS_HIP(hipGetLastError());
        return QCS_OK;
    }
    // sets of capacity cap each: records, 3 rows per ROI, chosen counts, device counts
    const size_t rw[3] = {(size_t)QCS_ROI_WORDS(7) * ncap, (size_t)QCS_ROI_WORDS(9) * ncap, (size_t)QCS_ROI_WORDS(11) * ncap};
    const size_t rec_bytes = ((rw[0] + rw[1] + rw[2]) * sizeof(uint16_t) + 255) & ~(size_t)255;
    const int ntiles = scan_tile_count((size_t)ncap);
    int rc;
    if ((rc = ctx->mx_sets.ensure(rec_bytes + sizeof(int) * (3 * (size_t)ncap + 16))) ||
        (rc = ctx->mx_rows.ensure(sizeof(float) * QCS_REC_FLOATS * 9 * (size_t)ncap + sizeof(int) * (2 * (size_t)ncap + ntiles + 1))))
        return rc;
    uint16_t *recs = ctx->mx_sets.as<uint16_t>();
    int *sel = reinterpret_cast<int *>(ctx->mx_sets.as<char>() + rec_bytes);
    int *cnts = sel + 3 * (size_t)ncap;
    float *rows = ctx->mx_rows.as<float>();
    int *rcnt = reinterpret_cast<int *>(rows + (size_t)QCS_REC_FLOATS * 9 * ncap);
    int *roff = rcnt + ncap, *rts = roff + ncap;
    qcs_roi_set sets[3];
    size_t ro = 0;
    for (int k = 0; k < 3; k++) {
        sets[k].roi_size = 7 + 2 * k;
        sets[k].n = cap;
        sets[k].d_rois = recs + ro;
        sets[k].d_loc = rows + (size_t)QCS_REC_FLOATS * 3 * ncap * k;
        sets[k].d_sel = sel + (size_t)ncap * k;
        sets[k].d_n = cnts + k;
        ro += rw[k];
    }
    int *cand[3], *info = nullptr;
    hipEvent_t ev;
    if ((rc = timer_begin(ctx->timer, s, &ev))) return rc;
    if ((rc = extract_mixed_launch(ctx, d_frames, nframes, start_frame, d_start, cap, sets, cand, &info, s))) return rc;
    if ((rc = timer_end(ctx->timer, TM_EXTRACT, s, ev, nframes))) return rc;
    if ((rc = qcs_fit_mle_mixed_dev(ctx, sets, 3, kmax, s))) return rc;
    // rows in candidate order: per-candidate counts, their scan, then the gather
    fill_async(rcnt, ncap, 0, s);
    const dim3 g((ncap + 255) / 256), b(256);
    for (int k = 0; k < 3; k++)
        hipLaunchKernelGGL(frame_rows_count_kernel, g, b, 0, s, cand[k], sets[k].d_n, cap, sets[k].d_sel, rcnt);
    QCS_HIP(hipGetLastError());
    if ((rc = scan_exclusive(rcnt, ncap, roff, rts, s))) return rc;
    for (int k = 0; k < 3; k++)
        hipLaunchKernelGGL(frame_rows_gather_kernel, g, b, 0, s, cand[k], sets[k].d_n, cap, sets[k].d_sel, roff, sets[k].d_loc, out_cap,
                           d_out);
    hipLaunchKernelGGL(frame_rows_total_kernel, dim3(1), dim3(1), 0, s, rts, ntiles, out_cap, cnts, info, d_n_out);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

}  // namespace qcs

extern "C" int qcs_fit_warm_dev(qcs_ctx *ctx, qcs_net *net, const uint16_t *d_rois, int n, float *d_loc, int *d_sel,
                                long long *h_stats, void *stream)
{
    if (!ctx || !net || n < 0 || (n > 0 && (!d_rois || !d_loc))) return fail(QCS_ERR_ARG, "invalid argument");
    if (n == 0) return QCS_OK;
    return warm_launch(ctx, net, d_rois, n, d_loc, d_sel, h_stats, (hipStream_t)stream);
}

extern "C" int qcs_fit_plain_stats_dev(qcs_ctx *ctx, const uint16_t *d_rois, int n, int kmax, float *d_loc, int *d_sel,
                                       long long *h_stats, void *stream)
{
    if (!ctx || n < 0 || (n > 0 && (!d_rois || !d_loc)) || kmax < 1 || kmax > 3 || !h_stats)
        return fail(QCS_ERR_ARG, "invalid argument");
    for (int i = 0; i < 9; i++) h_stats[i] = 0;
    if (n == 0) return QCS_OK;
    hipStream_t s = (hipStream_t)stream;
    MeScratch m;
    int rc;
    if ((rc = me_scratch(ctx, n, false, true, m))) return rc;
    int *sel = d_sel ? d_sel : m.sel;
    fill_async(m.cnt, 64, 0, s);
    fill_async(m.info1, (size_t)12 * (((size_t)n + 3) & ~(size_t)3), 0, s);
    if ((rc = plain_path(ctx, kmax, d_rois, nullptr, n, nullptr, m.t1, m.t2, m.t3, m.l1, m.l2, m.l3, sel, m.cnt, m.list3, d_loc,
                         m.info1, m.info2, m.info3, m.kcnt, m.klist, s)))
        return rc;
    return stats_copy(ctx, m, n, sel, h_stats, s);
}

extern "C" int qcs_label_dev(qcs_ctx *ctx, const uint16_t *d_rois, int n, float *d_labels, int *d_sel, void *stream)
{
    if (!ctx || n < 0 || (n > 0 && (!d_rois || !d_labels))) return fail(QCS_ERR_ARG, "invalid argument");
    if (ctx->p.LocType != QCS_LOCTYPE_GS2D) return fail(QCS_ERR_ARG, "training labels are made with the 2D model only");
    if (n == 0) return QCS_OK;
    hipStream_t s = (hipStream_t)stream;
    MeScratch m;
    int rc;
    if ((rc = me_scratch(ctx, n, false, false, m))) return rc;
    if ((rc = ctx->cf_rows.ensure((size_t)n * 3 * QCS_REC_FLOATS * sizeof(float)))) return rc;
    int *sel = d_sel ? d_sel : m.sel;
    if ((rc = plain_path(ctx, 3, d_rois, nullptr, n, nullptr, m.t1, m.t2, m.t3, m.l1, m.l2, m.l3, sel, m.cnt, m.list3,
                         ctx->cf_rows.as<float>(), nullptr, nullptr, nullptr, m.kcnt, m.klist, s)))
        return rc;
    hipLaunchKernelGGL(label_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, sel, m.t1, m.t2, m.t3, d_labels);
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
    if (!(t->ftol >= 0.0f) || (t->active_set != 0 && t->active_set != 1)) return fail(QCS_ERR_ARG, "ftol must be >= 0 and active_set 0 or 1");
    if (t->pin_stop != 0 && t->pin_stop != 1) return fail(QCS_ERR_ARG, "pin_stop must be 0 or 1");
    ctx->tune = *t;
    return QCS_OK;
}
