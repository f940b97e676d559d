// cfq.hip -- persistent work-queue kernels of the converged multi-emitter fitter (gfx950).
//
// The one-pass kernel (cfit_kernel) keeps a wave busy until the slowest of its 64 fits ends.
// The iteration counts of multi-emitter fits are heavy-tailed (most converge in 5-10 BFGS steps,
// a few run to the cap), so most lanes of such a wave idle most of the time.  Here each lane runs
// its own fit as a small state machine and takes the next ROI from a device queue as soon as
// its fit ends:
//
//   prep pass (one lane per work item, uniform work): observed image, starting vector, loss and
//     gradient there, inverse-Fisher starting inverse Hessian -> one record per item in HBM;
//   persistent pass (as many single-wave work-groups as the chip holds at once): per tick, lanes
//     without work take items from the queue (one atomic per wave), lanes holding a fresh
//     gradient form the quasi-Newton direction (or finish), and every lane in a line search
//     evaluates loss and gradient at its trial point -- the one heavy, uniform step of the tick.
//     The gradient of an accepted trial is the gradient of the next iteration, so an iteration
//     costs one evaluation where the one-pass kernel spends two;
//   likelihood pass (one lane per work item): FP64 log-likelihood at the final parameters.
//
// A launch holds up to three sets of ROIs (sizes 7, 9, 11, each with its own queue): every wave
// drains the sets in order with its LDS sized for the largest.  Each step's arithmetic is that
// of the one-pass kernel, so both give bit-identical parameters, iteration counts and flags (the
// evaluation counts differ by the saved gradient evaluations).
#include <string.h>
#include <type_traits>
#include "cfit_dev.h"

namespace qcs {

// record of one work item: theta, loss, gradient, "identity start" flag, packed D0, padding
template <int KK> struct CQ {
    static constexpr int P = CP<KK>::P, NS = CP<KK>::NS;
    static constexpr int OL = P, OG = P + 1, OF = 2 * P + 1, OD = 2 * P + 2;
    static constexpr int W = (2 * P + 2 + NS + 3) & ~3;   // floats, a multiple of 4
};

enum { CQ_IDLE = 0, CQ_DIR = 1, CQ_LS = 2 };

template <int KK, int R>
__global__ void __launch_bounds__(64) cfq_prep_kernel(CfqSet S, CFitConst c)
{
    using Q = CQ<KK>;
    constexpr int P = Q::P, NS = Q::NS;
    // the start matrix is factored in registers and written to the record from there: LDS holds
    // only the observed image (the 16-parameter model's LDS matrix had left two work-groups per CU)
    __shared__ float qsh[R * R * 64];
    const int lane = threadIdx.x, j = blockIdx.x * 64 + lane;
    const int n = S.n_dev ? min(*S.n_dev, S.n) : S.n;   // a device count never exceeds the capacity
    if (j >= n) return;   // no cross-lane work here
    const int roi = S.idx ? S.idx[j] : j;
    float *qs = qsh + lane;
    cf_load_q<R>(S.rois, roi, c, qs);
    float th[P], g[P];
    cf_init<KK, R>(qs, S.init ? S.init + (size_t)roi * S.istride : nullptr, S.init ? S.init_mode : CF_INIT_HEUR, c.sigma0, th);
    cf_project<KK, R>(th);
    const float L = cf_lossgrad<KK, R, true>(qs, th, g);
    float Dr[NS];
    const bool first = !cf_fisher_inverse_regs<KK, R>(th, Dr);
    if (first)
        cf_sfor<0, P>([&](auto IC) {
            constexpr int i = decltype(IC)::value;
            cf_sfor<i, P>([&](auto JC) {
                constexpr int jj = decltype(JC)::value;
                Dr[tri(i, jj, P)] = (i == jj) ? 1.0f : 0.0f;
            });
        });
    float4 *rec = reinterpret_cast<float4 *>(S.rec + (size_t)j * Q::W);
#pragma unroll
    for (int v = 0; v < Q::W / 4; v++) {
        float e[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = 4 * v + u;
            e[u] = i < P ? th[i < P ? i : 0]
                 : i == Q::OL ? L
                 : i < Q::OF ? g[(i > Q::OL && i < Q::OF) ? i - Q::OG : 0]
                 : i == Q::OF ? (first ? 1.0f : 0.0f)
                 : i < Q::OD + NS ? Dr[(i >= Q::OD && i < Q::OD + NS) ? i - Q::OD : 0] : 0.0f;
        }
        rec[v] = make_float4(e[0], e[1], e[2], e[3]);
    }
}

// Storage of the persistent kernel for a model key and largest ROI size, chosen so that more
// single-wave work-groups fit in a CU: QU16 keeps the observed image as raw counts (2 B per pixel
// in LDS, converted on use), the first NREG packed inverse-Hessian entries live in registers, and
// WAVES is the register budget's target in waves per SIMD.  The arithmetic is the same in every
// layout.  The 2D three-emitter 7x7 kernel gains from it (5 -> 8 work-groups per CU, 4.78 ->
// 4.25 ms on 2^20 ROIs), and so do the 2D mixed-size launches (high-density frames 9% faster);
// the elliptical and the other 7x7 kernels were measured at higher occupancy and did not (the
// register-held entries cost moves, the converted counts cost 5 instructions per pixel).
template <int KK, int RMAX> struct CqLayout {
    static constexpr bool QU16 = false;
    static constexpr int NREG = 0, WAVES = 1;
};
template <> struct CqLayout<3, 7> {   // 24.5 + 55 floats of LDS per lane: 8 work-groups per CU
    static constexpr bool QU16 = true;
    static constexpr int NREG = 11, WAVES = 2;
};
// mixed-size launches (sets of 7, 9 and 11 pixels, LDS sized for 11): raw counts
template <> struct CqLayout<1, 11> {   // 60.5 + 15 floats: 8 per CU
    static constexpr bool QU16 = true;
    static constexpr int NREG = 0, WAVES = 2;
};
template <> struct CqLayout<2, 11> {   // 60.5 + 36 floats: 6 per CU
    static constexpr bool QU16 = true;
    static constexpr int NREG = 0, WAVES = 2;
};
template <> struct CqLayout<3, 11> {   // 60.5 + 66 floats: 5 per CU
    static constexpr bool QU16 = true;
    static constexpr int NREG = 0, WAVES = 2;
};
template <class LY> using CqQ = typename std::conditional<LY::QU16, uint16_t, float>::type;

constexpr int CQ_CHUNK = 64;   // most items a wave takes from the queue with one atomic

template <int KK, int R, class LY>
__device__ __forceinline__ void cfq_run(const CfqSet &S, int *head, const CFitConst &c, CqQ<LY> *qsh, float *Dsh)
{
    using Q = CQ<KK>;
    constexpr int P = Q::P;
    const int lane = threadIdx.x;
    CqQ<LY> *qs = qsh + lane;
    DSplit<LY::NREG> D;
    D.p = Dsh + lane;
    const auto qa = [&]() {
        if constexpr (LY::QU16) return QLds16{qs, c.offset, c.kadc, c.inv_qe};
        else return QLds{qs};
    }();
    const int n = S.n_dev ? min(*S.n_dev, S.n) : S.n;   // a device count never exceeds the capacity
    int st = CQ_IDLE, roi = 0, iters = 0, flags = 0, evals = 0, t = 0, pin_run = 0;
    bool first = false, dry = n <= 0;
    float th[P], g[P], du[P], L = 0.0f, slope = 0.0f, alpha = 1.0f, dmax = 0.0f;
#pragma unroll
    for (int i = 0; i < P; i++) th[i] = g[i] = du[i] = 0.0f;

    auto finish = [&]() {
        float *o = S.theta + (size_t)roi * S.ostride;
#pragma unroll
        for (int i = 0; i < P; i++) o[i] = th[i];
        if (S.info) {
            int *inf = S.info + (size_t)roi * 4;
            inf[0] = iters;
            inf[1] = flags;
            inf[2] = evals;
            inf[3] = 0;
        }
        st = CQ_IDLE;
    };
    // Lanes without work take the next items of the queue.  The wave keeps a chunk of taken but
    // unstarted items [cb, cb + cl) (wave-uniform) and refills it with one atomic when the idle
    // lanes need more than it holds: up to CQ_CHUNK items, but no more than the wave's share of
    // what the queue held at its last refill, so near the end no wave holds back work that others
    // could run.  Same-address atomics serialise at the L2; one per refill of a few lanes made
    // the queue head a hot spot.
    int cb = 0, cl = 0, seen = 0;
    bool qdone = false;
    auto refill = [&]() {
        const bool want = st == CQ_IDLE;
        const unsigned long long m = __ballot(want);
        if (m == 0ull || dry) return;   // wave-uniform
        const int cnt = __popcll(m), leader = __ffsll((unsigned long long)m) - 1;
        const int rank = __popcll(m & ((1ull << lane) - 1ull));
        int j;
        if (cnt <= cl) {
            j = cb + rank;
            cb += cnt;
            cl -= cnt;
        } else {
            const int need = cnt - cl;
            const int grab = max(need, min(CQ_CHUNK, (n - seen) / (int)gridDim.x));
            int base = 0;
            if (lane == leader) base = atomicAdd(head, grab);
            base = __shfl(base, leader, 64);
            j = rank < cl ? cb + rank : base + (rank - cl);
            const int got = max(0, min(grab, n - base));   // items of the grab inside the queue
            const int used = min(need, got);
            cb = base + used;
            cl = got - used;
            seen = min(base + grab, n);
            qdone = base + grab >= n;
        }
        if (qdone && cl == 0) dry = true;   // the queue holds nothing more for this wave
        if (!want) return;
        if (j >= n) return;
        roi = S.idx ? S.idx[j] : j;
        if constexpr (LY::QU16) cf_load_q16<R>(S.rois, roi, qs);
        else cf_load_q<R>(S.rois, roi, c, qs);
        const float4 *rec = reinterpret_cast<const float4 *>(S.rec + (size_t)j * Q::W);
#pragma unroll
        for (int v = 0; v < Q::W / 4; v++) {
            const float4 x = rec[v];
            const float e[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int i = 4 * v + u;
                if (i < P) th[i < P ? i : 0] = e[u];
                else if (i == Q::OL) L = e[u];
                else if (i < Q::OF) g[(i > Q::OL && i < Q::OF) ? i - Q::OG : 0] = e[u];
                else if (i == Q::OF) first = e[u] != 0.0f;
                else if (i < Q::OD + Q::NS) D[(i >= Q::OD && i < Q::OD + Q::NS) ? i - Q::OD : 0] = e[u];
            }
        }
        iters = 0;
        flags = 0;
        evals = 1;
        pin_run = 0;
        st = CQ_DIR;
    };
    // quasi-Newton direction from a fresh gradient; converged fits end here
    auto direction = [&]() {
        if (st != CQ_DIR) return;
        if (cf_direction<KK, R>(th, g, D, first, c.tol, c.active_set != 0, du, slope, dmax)) {
            flags |= CF_CONVERGED;
            finish();
            return;
        }
        alpha = 1.0f;
        t = 0;
        st = CQ_LS;
    };

    while (true) {
        refill();
        direction();
        refill();      // lanes whose fit converged just now
        direction();
        if (dry && __all(st == CQ_IDLE)) break;
        if (st != CQ_LS) continue;
        float tn[P], gt[P];
#pragma unroll
        for (int i = 0; i < P; i++) tn[i] = fmaf(alpha, du[i] * cf_scale<KK>(i), th[i]);
        cf_project<KK, R>(tn);
        const float Lt = cf_lossgrad_q<KK, R, true>(qa, tn, gt);
        evals++;
        if (Lt == Lt && Lt < L && Lt <= fmaf(1e-4f * alpha, slope, L)) {
            // accepted: BFGS update of the inverse Hessian with s = step, y = gradient change
            iters++;
            cf_bfgs_update<KK>(D, first, th, tn, g, gt);
#pragma unroll
            for (int i = 0; i < P; i++) {
                th[i] = tn[i];
                g[i] = gt[i];
            }
            const float Lold = L;
            L = Lt;
            if constexpr (CM<KK>::AS && CM<KK>::K == 3)
                if (c.pin_stop) pin_run = cf_at_bound<KK, R>(th) ? pin_run + 1 : 0;
            if (Lold - L < c.ftol) {   // the likelihood has stopped moving
                flags |= CF_CONVERGED | CF_FLAT;
                finish();
            } else if (pin_run >= CF_PIN_RUN) {
                flags |= CF_PINNED;
                finish();
            } else if (iters >= c.max_iter) {
                flags |= CF_MAXITER;
                finish();
            } else {
                st = CQ_DIR;
            }
        } else if (t == 0 && !(Lt < L) && dmax < 20.0f * c.tol) {
            // the full step is tiny and the FP32 loss cannot resolve it: take it and stop
            iters++;
#pragma unroll
            for (int i = 0; i < P; i++) th[i] = tn[i];
            flags |= CF_CONVERGED;
            finish();
        } else {
            alpha *= 0.3f;
            if (++t >= 16) {   // no decrease in 16 backtracks
                iters++;
                flags |= CF_STALLED;
                finish();
            }
        }
    }
}

template <int KK, int RMAX>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(CqLayout<KK, RMAX>::WAVES))) cfq_kernel(CfqArgs a)
{
    using LY = CqLayout<KK, RMAX>;
    constexpr int NDL = CP<KK>::NS - LY::NREG;
    __shared__ CqQ<LY> qsh[RMAX * RMAX * 64];
    __shared__ float Dsh[(NDL > 0 ? NDL : 1) * 64];
    for (int si = 0; si < a.nsets; si++) {
        const int R = a.set[si].R;
        if constexpr (RMAX >= 11) {
            if (R == 11) cfq_run<KK, 11, LY>(a.set[si], a.heads + si, a.c, qsh, Dsh);
        }
        if constexpr (RMAX >= 9) {
            if (R == 9) cfq_run<KK, 9, LY>(a.set[si], a.heads + si, a.c, qsh, Dsh);
        }
        if (R == 7) cfq_run<KK, 7, LY>(a.set[si], a.heads + si, a.c, qsh, Dsh);
    }
}

template <int KK, int R>
__global__ void __launch_bounds__(64) cfq_ll_kernel(CfqSet S, CFitConst c)
{
    __shared__ float qsh[R * R * 64];
    const int lane = threadIdx.x, j = blockIdx.x * 64 + lane;
    const int n = S.n_dev ? min(*S.n_dev, S.n) : S.n;   // a device count never exceeds the capacity
    if (j >= n) return;
    const int roi = S.idx ? S.idx[j] : j;
    float *qs = qsh + lane;
    cf_load_q<R>(S.rois, roi, c, qs);
    float th[CP<KK>::P];
#pragma unroll
    for (int i = 0; i < CP<KK>::P; i++) th[i] = S.theta[(size_t)roi * S.ostride + i];
    S.ll[roi] = cf_loglik64<KK, R>(qs, th);
}

// ---- host -----------------------------------------------------------------------------------

static int cq_rec_w(int KK)
{
    switch (KK) {
    case 1: return CQ<1>::W;
    case 2: return CQ<2>::W;
    case 3: return CQ<3>::W;
    case 11: return CQ<11>::W;
    case 12: return CQ<12>::W;
    default: return CQ<13>::W;
    }
}

// resident work-groups per CU of kernel f
static int cq_occupancy(const void *f)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64, 0) != hipSuccess || nb < 1) nb = 1;
    return nb;
}

// resident persistent work-groups per CU for model key KK and largest ROI size RMAX (cached)
static int cq_waves_per_cu(int KK, int RMAX)
{
    static int cache[2][3][3] = {{{0}}};
    const int ri = RMAX == 7 ? 0 : (RMAX == 9 ? 1 : 2);
    const int as = KK > CF_ELLIPTIC ? 1 : 0, K = as ? KK - CF_ELLIPTIC : KK;
    int &v = cache[as][K - 1][ri];
    if (v == 0) {
        const void *f = nullptr;
        CF_DISPATCH_KK(KK, CF_DISPATCH_R(RMAX, f = reinterpret_cast<const void *>(&cfq_kernel<KT, RT>)))
        v = cq_occupancy(f);
    }
    return v;
}

int cfq_launch(qcs_ctx *ctx, int K, const CfqSet *sets_in, int nsets, hipStream_t s)
{
    if (K < 1 || K > 3) return fail(QCS_ERR_ARG, "emitter count %d not in 1..3", K);
    if (nsets < 1 || nsets > 3) return fail(QCS_ERR_ARG, "1..3 ROI sets per launch");
    CfqArgs a;
    memset(&a, 0, sizeof(a));
    a.c = cfit_const(ctx->p, ctx->tune, K, CF_INIT_HEUR);
    const int KK = cf_kk(ctx->p, K);
    const int W = cq_rec_w(KK);
    size_t total = 0;
    int rmax = 0;
    long long items = 0;
    for (int i = 0; i < nsets; i++) {
        const CfqSet &x = sets_in[i];
        if (x.R != 7 && x.R != 9 && x.R != 11) return fail(QCS_ERR_ARG, "converged fitting supports ROISize 7, 9, 11 (got %d)", x.R);
        if (x.n < 0 || (x.n > 0 && (!x.rois || !x.theta))) return fail(QCS_ERR_ARG, "invalid ROI set");
        if (x.init_mode < CF_INIT_HEUR || x.init_mode > CF_INIT_ADD) return fail(QCS_ERR_ARG, "invalid start mode");
        if (x.init_mode == CF_INIT_ADD && K < 2) return fail(QCS_ERR_ARG, "adding an emitter needs K >= 2");
        if (x.init_mode != CF_INIT_HEUR && x.n > 0 && !x.init) return fail(QCS_ERR_ARG, "this start needs an initial vector");
        total += (size_t)x.n * W;
        items += x.n;
        rmax = x.R > rmax ? x.R : rmax;
    }
    if (items == 0) return QCS_OK;
    int rc;
    if ((rc = ctx->cf_queue.ensure(256 + total * sizeof(float)))) return rc;
    int *heads = ctx->cf_queue.as<int>();
    float *rec = reinterpret_cast<float *>(ctx->cf_queue.as<char>() + 256);
    // largest ROIs first: the longest fits start early and the tail stays short
    int order[3] = {0, 1, 2};
    for (int i = 0; i < nsets; i++)
        for (int j2 = i + 1; j2 < nsets; j2++)
            if (sets_in[order[j2]].R > sets_in[order[i]].R) { const int x = order[i]; order[i] = order[j2]; order[j2] = x; }
    size_t off = 0;
    a.nsets = 0;
    for (int i = 0; i < nsets; i++) {
        const CfqSet &x = sets_in[order[i]];
        if (x.n == 0) continue;
        CfqSet &y = a.set[a.nsets++];
        y = x;
        y.rec = rec + off;
        off += (size_t)x.n * W;
    }
    a.heads = heads;
    fill_async(heads, 4, 0, s);
    for (int i = 0; i < a.nsets; i++) {
        const CfqSet &x = a.set[i];
        const dim3 grid((x.n + 63) / 64), block(64);
        CF_DISPATCH_KK(KK, CF_DISPATCH_R(x.R, hipLaunchKernelGGL((cfq_prep_kernel<KT, RT>), grid, block, 0, s, x, a.c)))
    }
    QCS_HIP(hipGetLastError());
    int cus = 0;
    QCS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    int wpc = cq_waves_per_cu(KK, rmax);
    if (ctx->tune.waves_per_cu > 0 && ctx->tune.waves_per_cu < wpc) wpc = ctx->tune.waves_per_cu;
    long long waves = (items + 63) / 64;
    if (waves > (long long)cus * wpc) waves = (long long)cus * wpc;
    const dim3 pgrid((unsigned)waves), block(64);
    CF_DISPATCH_KK(KK, CF_DISPATCH_R(rmax, hipLaunchKernelGGL((cfq_kernel<KT, RT>), pgrid, block, 0, s, a)))
    QCS_HIP(hipGetLastError());
    for (int i = 0; i < a.nsets; i++) {
        const CfqSet &x = a.set[i];
        if (!x.ll) continue;
        const dim3 grid((x.n + 63) / 64);
        CF_DISPATCH_KK(KK, CF_DISPATCH_R(x.R, hipLaunchKernelGGL((cfq_ll_kernel<KT, RT>), grid, block, 0, s, x, a.c)))
    }
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

}  // namespace qcs
