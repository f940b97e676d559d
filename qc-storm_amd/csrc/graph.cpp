// graph.cpp -- per-batch HIP graphs of the converged multi-emitter paths.
//
// The plain and warm-started multi-emitter paths run entirely on device-side counts (work lists
// built by the kernels, counters cleared by fill kernels, no host synchronisation), so one call can
// be captured into a HIP graph and replayed per batch: the launch sequence (up to ~20 kernels)
// then costs one graph launch.  The graph owns a private context, so its scratch buffers keep the
// addresses the captured kernels use however the caller's context is used meanwhile.
#include <new>
#include "qcs_internal.h"

using namespace qcs;

struct qcs_graph {
    qcs_ctx *ctx = nullptr;       // private: scratch addresses are baked into the graph
    hipStream_t cap = nullptr;    // capture stream
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
};

namespace {

int run_once(qcs_graph *G, qcs_net *net, const uint16_t *d_rois, int n, int kmax, float *d_loc, int *d_sel)
{
    if (net) return qcs_fit_warm_dev(G->ctx, net, d_rois, n, d_loc, d_sel, nullptr, G->cap);
    return qcs_fit_mle_dev(G->ctx, d_rois, n, kmax, d_loc, d_sel, G->cap);
}

}  // namespace

extern "C" int qcs_graph_destroy(qcs_graph *G)
{
    if (!G) return QCS_OK;
    if (G->ge) (void)hipGraphExecDestroy(G->ge);
    if (G->g) (void)hipGraphDestroy(G->g);
    if (G->cap) (void)hipStreamDestroy(G->cap);
    if (G->ctx) (void)qcs_ctx_destroy(G->ctx);
    delete G;
    return QCS_OK;
}

extern "C" int qcs_graph_create_mle(qcs_ctx *ctx, qcs_net *net, const uint16_t *d_rois, int n, int kmax, float *d_loc, int *d_sel,
                                    void *stream, qcs_graph **out)
{
    if (!ctx || !out || n <= 0 || !d_rois || !d_loc || kmax < 1 || kmax > 3) return fail(QCS_ERR_ARG, "invalid argument");
    *out = nullptr;
    qcs_graph *G = new (std::nothrow) qcs_graph();
    if (!G) return fail(QCS_ERR_NOMEM, "graph allocation failed");
    int rc = qcs_ctx_create(&ctx->p, &G->ctx);
    if (rc) {
        qcs_graph_destroy(G);
        return rc;
    }
    G->ctx->tune = ctx->tune;
    hipError_t e = hipStreamCreateWithFlags(&G->cap, hipStreamNonBlocking);
    if (e != hipSuccess) {
        qcs_graph_destroy(G);
        return hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    }
    // the inputs must be ready on the caller's stream; one eager run sizes every scratch buffer
    if ((e = hipStreamSynchronize((hipStream_t)stream)) != hipSuccess) {
        qcs_graph_destroy(G);
        return hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
    }
    if ((rc = run_once(G, net, d_rois, n, kmax, d_loc, d_sel)) || (e = hipStreamSynchronize(G->cap)) != hipSuccess) {
        qcs_graph_destroy(G);
        return rc ? rc : hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
    }
    if ((e = hipStreamBeginCapture(G->cap, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
        qcs_graph_destroy(G);
        return hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
    }
    rc = run_once(G, net, d_rois, n, kmax, d_loc, d_sel);
    e = hipStreamEndCapture(G->cap, &G->g);
    if (rc || e != hipSuccess) {
        qcs_graph_destroy(G);
        return rc ? rc : hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    }
    if ((e = hipGraphInstantiate(&G->ge, G->g, nullptr, nullptr, 0)) != hipSuccess) {
        qcs_graph_destroy(G);
        return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    *out = G;
    return QCS_OK;
}

extern "C" int qcs_graph_launch(qcs_graph *G, void *stream)
{
    if (!G || !G->ge) return fail(QCS_ERR_ARG, "invalid graph");
    QCS_HIP(hipGraphLaunch(G->ge, (hipStream_t)stream));
    return QCS_OK;
}
