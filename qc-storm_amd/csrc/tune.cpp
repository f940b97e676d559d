// tune.cpp -- autotuner of the converged multi-emitter paths (no counterpart in the reference).
//
// On a sample of ROIs the paths are first run with a fully converged reference setting (caps of
// 100 iterations, no loss-decrease stop); every candidate setting (iteration caps, loss-decrease
// tolerance, persistent waves per CU, persistent vs one-pass kernels and -- with a network -- the
// warm start's confidence threshold) is then timed with HIP events on the caller's stream and
// compared with the reference: the fraction of ROIs whose emitter count agrees, and the 99th
// percentile of the position differences over those ROIs.  The fastest candidate within the
// caller's tolerances becomes the context's tuning.
#include <math.h>
#include <string.h>
#include <algorithm>
#include <vector>
#include "qcs_internal.h"

using namespace qcs;

namespace {

struct Run {
    std::vector<float> loc;
    std::vector<int> sel;
};

int run_path(qcs_ctx *ctx, qcs_net *net, const uint16_t *d_rois, int n, float *d_loc, int *d_sel, hipStream_t s)
{
    if (net) return qcs_fit_warm_dev(ctx, net, d_rois, n, d_loc, d_sel, nullptr, s);
    return qcs_fit_mle_dev(ctx, d_rois, n, 3, d_loc, d_sel, s);
}

int fetch(const float *d_loc, const int *d_sel, int n, hipStream_t s, Run &r)
{
    r.loc.resize((size_t)n * 3 * QCS_REC_FLOATS);
    r.sel.resize((size_t)n);
    QCS_HIP(hipStreamSynchronize(s));
    QCS_HIP(hipMemcpy(r.loc.data(), d_loc, r.loc.size() * sizeof(float), hipMemcpyDeviceToHost));
    QCS_HIP(hipMemcpy(r.sel.data(), d_sel, r.sel.size() * sizeof(int), hipMemcpyDeviceToHost));
    return QCS_OK;
}

// agreement of a run with the reference: count agreement and p99 of max(|dx|, |dy|) (px) over the
// emitters of the ROIs whose counts agree, emitters paired by the permutation with the smallest
// summed squared distance (the fits may order them differently)
void compare(const Run &ref, const Run &r, float &agree, float &dx_p99)
{
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    const size_t n = ref.sel.size();
    size_t same = 0;
    std::vector<float> d;
    d.reserve(n);
    for (size_t j = 0; j < n; j++) {
        const int K = ref.sel[j];
        if (K != r.sel[j]) continue;
        same++;
        if (K < 1 || K > 3) continue;
        const float *a = &ref.loc[j * 3 * QCS_REC_FLOATS], *b = &r.loc[j * 3 * QCS_REC_FLOATS];
        auto ex = [&](const float *row, int k, int w) { return row[k * QCS_REC_FLOATS + w]; };
        double best = 1e300;
        int bp = 0;
        for (int pi = 0; pi < 6; pi++) {
            bool valid = true;
            for (int k = 0; k < K; k++) valid = valid && perms[pi][k] < K;
            if (!valid) continue;
            double ss = 0.0;
            for (int k = 0; k < K; k++) {
                const double dx = ex(a, k, QCS_POS_XPOS) - ex(b, perms[pi][k], QCS_POS_XPOS);
                const double dy = ex(a, k, QCS_POS_YPOS) - ex(b, perms[pi][k], QCS_POS_YPOS);
                ss += dx * dx + dy * dy;
            }
            if (ss < best) {
                best = ss;
                bp = pi;
            }
        }
        for (int k = 0; k < K; k++) {
            const float dd = std::max(fabsf(ex(a, k, QCS_POS_XPOS) - ex(b, perms[bp][k], QCS_POS_XPOS)),
                                      fabsf(ex(a, k, QCS_POS_YPOS) - ex(b, perms[bp][k], QCS_POS_YPOS)));
            d.push_back(dd == dd ? dd : 1e30f);
        }
    }
    agree = n ? (float)same / (float)n : 1.0f;
    if (d.empty()) {
        dx_p99 = 0.0f;
        return;
    }
    const size_t q = std::min(d.size() - 1, (size_t)floor(0.99 * (double)(d.size() - 1) + 0.5));
    std::nth_element(d.begin(), d.begin() + q, d.end());
    dx_p99 = d[q];
}

}  // namespace

extern "C" int qcs_autotune_mle_dev(qcs_ctx *ctx, qcs_net *net, const uint16_t *d_rois, int n, float *d_loc, int *d_sel,
                                    const qcs_autotune_opts *opts, qcs_autotune_entry *table, int table_cap, int *n_entries,
                                    void *stream)
{
    if (!ctx || !d_rois || n <= 0 || !d_loc || !d_sel || !opts || (table_cap > 0 && !table))
        return fail(QCS_ERR_ARG, "invalid argument");
    if (opts->trials < 1 || opts->trials > 100 || !(opts->min_select_agree >= 0.0f && opts->min_select_agree <= 1.0f) ||
        !(opts->max_dx_p99 >= 0.0f))
        return fail(QCS_ERR_ARG, "invalid autotune options");
    if (net && ctx->p.ROISize != 7) return fail(QCS_ERR_ARG, "the warm start takes 7x7 ROIs");
    if (net && ctx->p.LocType != QCS_LOCTYPE_GS2D) return fail(QCS_ERR_ARG, "the warm start supports the 2D model only");
    // the elliptical model's fits fix their own caps and stopping rule: only the kernel choice is tuned
    const bool elliptical = ctx->p.LocType == QCS_LOCTYPE_AS3D;
    hipStream_t s = (hipStream_t)stream;
    const qcs_tuning base = ctx->tune;
    int rc = QCS_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    QCS_HIP(hipEventCreate(&e0));
    QCS_HIP(hipEventCreate(&e1));

    // reference: the plain path run to convergence
    qcs_tuning ref_t = base;
    ref_t.max_iter[0] = ref_t.max_iter[1] = ref_t.max_iter[2] = 100;
    ref_t.ftol = 0.0f;
    ref_t.queue = 1;
    ref_t.waves_per_cu = 0;
    Run ref;
    ctx->tune = ref_t;
    if ((rc = run_path(ctx, nullptr, d_rois, n, d_loc, d_sel, s)) || (rc = fetch(d_loc, d_sel, n, s, ref))) {
        ctx->tune = base;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc;
    }

    // candidates: plain settings first, then (with a network) the warm path on the best plain one
    std::vector<qcs_tuning> cand;
    const int caps[][3] = {{100, 60, 60}, {100, 40, 40}, {100, 30, 30}, {100, 20, 25}};
    const float ftols[] = {0.0f, 1e-3f, 1e-2f};
    for (int ci = 0; ci < 4; ci++)
        for (int fi = 0; fi < 3; fi++) {
            if (elliptical && (ci > 0 || fi > 0)) continue;
            qcs_tuning t = base;
            for (int k = 0; k < 3; k++) t.max_iter[k] = caps[ci][k];
            t.ftol = ftols[fi];
            t.queue = 1;
            t.waves_per_cu = 0;
            cand.push_back(t);
        }
    {
        qcs_tuning t = base;   // persistent grid sizes and the one-pass kernels at the default caps
        t.max_iter[0] = 100;
        t.max_iter[1] = t.max_iter[2] = 40;
        t.ftol = 1e-3f;
        for (int w : {1, 2, 4}) {
            t.queue = 1;
            t.waves_per_cu = w;
            cand.push_back(t);
        }
        t.queue = 0;
        t.waves_per_cu = 0;
        cand.push_back(t);
    }
    struct Res {
        qcs_tuning t;
        float ms, agree, dx;
        int warm, ok;
    };
    std::vector<Res> res;
    auto measure = [&](const qcs_tuning &t, qcs_net *nt, Res &out) -> int {
        ctx->tune = t;
        int r2;
        if ((r2 = run_path(ctx, nt, d_rois, n, d_loc, d_sel, s))) return r2;   // warm-up
        QCS_HIP(hipEventRecord(e0, s));
        for (int i = 0; i < opts->trials; i++)
            if ((r2 = run_path(ctx, nt, d_rois, n, d_loc, d_sel, s))) return r2;
        QCS_HIP(hipEventRecord(e1, s));
        QCS_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        QCS_HIP(hipEventElapsedTime(&ms, e0, e1));
        Run r;
        if ((r2 = fetch(d_loc, d_sel, n, s, r))) return r2;
        out.t = t;
        out.ms = ms / (float)opts->trials;
        compare(ref, r, out.agree, out.dx);
        out.warm = nt != nullptr;
        out.ok = out.agree >= opts->min_select_agree && out.dx <= opts->max_dx_p99;
        return QCS_OK;
    };
    for (const auto &t : cand) {
        Res r;
        if ((rc = measure(t, nullptr, r))) break;
        res.push_back(r);
    }
    if (!rc && net) {
        int best = -1;
        for (size_t i = 0; i < res.size(); i++)
            if (res[i].ok && (best < 0 || res[i].ms < res[best].ms)) best = (int)i;
        qcs_tuning t = best >= 0 ? res[best].t : base;
        for (float pm : {0.7f, 0.8f, 0.9f, 0.95f}) {
            t.warm_pmin = pm;
            Res r;
            if ((rc = measure(t, net, r))) break;
            res.push_back(r);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        ctx->tune = base;
        return rc;
    }
    // the fastest plain candidate within the tolerances sets the fitting knobs; the fastest warm
    // candidate within them sets warm_pmin
    int bp = -1, bw = -1;
    for (size_t i = 0; i < res.size(); i++) {
        if (!res[i].ok) continue;
        int &b = res[i].warm ? bw : bp;
        if (b < 0 || res[i].ms < res[b].ms) b = (int)i;
    }
    qcs_tuning chosen = bp >= 0 ? res[bp].t : base;
    if (bw >= 0) chosen.warm_pmin = res[bw].t.warm_pmin;
    else chosen.warm_pmin = base.warm_pmin;
    ctx->tune = chosen;
    const int m = std::min((int)res.size(), table_cap);
    for (int i = 0; i < m; i++) {
        table[i].tuning = res[i].t;
        table[i].ms = res[i].ms;
        table[i].select_agree = res[i].agree;
        table[i].dx_p99 = res[i].dx;
        table[i].warm = res[i].warm;
        table[i].within = res[i].ok;
        table[i].chosen = (int)i == bp || (int)i == bw;
    }
    if (n_entries) *n_entries = (int)res.size();
    return QCS_OK;
}
