// graph.cpp -- per-batch HIP graphs of the converged multi-emitter paths.
//
// The plain and warm-started multi-emitter paths and the frames -> rows path
// (qcs_loc_mle_frames_dev: mixed-size extraction, converged fits, frame-sorted rows) run entirely
// on device-side counts (work lists built by the kernels, counters cleared by fill kernels, no
// host synchronisation), so one call can be captured into a HIP graph and replayed per batch: the
// launch sequence (up to ~20 kernels, about 60 for a frame batch) then costs one graph launch.  The graph owns a private context, so its scratch buffers keep the
// addresses the captured kernels use however the caller's context is used meanwhile.
#include <new>
#include "qcs_internal.h"

namespace qcs {
int loc_mle_frames(qcs_ctx *ctx, const uint16_t *d_frames, int nframes, int start_frame, const int *d_start, int cap, int kmax,
                   float *d_out, int out_cap, int *d_n_out, hipStream_t s);
}

using namespace qcs;

struct qcs_graph {
    qcs_ctx *ctx = nullptr;       // private: scratch addresses are baked into the graph
    hipStream_t cap = nullptr;    // capture stream
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
};

namespace {

// the captured call: one multi-emitter batch over ROIs, or a frame batch through the whole
// frames -> rows path
struct GraphCall {
    qcs_net *net = nullptr;
    const uint16_t *d_rois = nullptr;
    int n = 0, kmax = 3;
    float *d_loc = nullptr;
    int *d_sel = nullptr;
    const uint16_t *d_frames = nullptr;
    int nframes = 0, start_frame = 1, cap = 0, out_cap = 0;
    int *d_n_out = nullptr;
    const int *d_start = nullptr;
};

int run_once(qcs_graph *G, const GraphCall &a)
{
    if (a.d_frames)
        return loc_mle_frames(G->ctx, a.d_frames, a.nframes, a.start_frame, a.d_start, a.cap, a.kmax, a.d_loc, a.out_cap, a.d_n_out,
                              G->cap);
    if (a.net) return qcs_fit_warm_dev(G->ctx, a.net, a.d_rois, a.n, a.d_loc, a.d_sel, nullptr, G->cap);
    return qcs_fit_mle_dev(G->ctx, a.d_rois, a.n, a.kmax, a.d_loc, a.d_sel, G->cap);
}

// private context, capture stream, one eager run (sizes every scratch buffer), capture, instantiate
int build_graph(qcs_ctx *ctx, const GraphCall &a, void *stream, qcs_graph **out)
{
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
    if ((rc = run_once(G, a)) || (e = hipStreamSynchronize(G->cap)) != hipSuccess) {
        qcs_graph_destroy(G);
        return rc ? rc : hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
    }
    if ((e = hipStreamBeginCapture(G->cap, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
        qcs_graph_destroy(G);
        return hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
    }
    rc = run_once(G, a);
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
    GraphCall a;
    a.net = net;
    a.d_rois = d_rois;
    a.n = n;
    a.kmax = kmax;
    a.d_loc = d_loc;
    a.d_sel = d_sel;
    return build_graph(ctx, a, stream, out);
}

extern "C" int qcs_graph_create_frames(qcs_ctx *ctx, const uint16_t *d_frames, int nframes, const int *d_start_frame, int cap,
                                       int kmax, float *d_out, int out_cap, int *d_n_out, void *stream, qcs_graph **out)
{
    if (!ctx || !out || !d_frames || nframes <= 0 || !d_start_frame || cap <= 0 || kmax < 1 || kmax > 3 || !d_out ||
        out_cap <= 0 || !d_n_out)
        return fail(QCS_ERR_ARG, "invalid argument");
    GraphCall a;
    a.d_frames = d_frames;
    a.nframes = nframes;
    a.d_start = d_start_frame;
    a.cap = cap;
    a.kmax = kmax;
    a.d_loc = d_out;
    a.out_cap = out_cap;
    a.d_n_out = d_n_out;
    return build_graph(ctx, a, stream, out);
}

extern "C" int qcs_graph_launch(qcs_graph *G, void *stream)
{
    if (!G || !G->ge) return fail(QCS_ERR_ARG, "invalid graph");
    QCS_HIP(hipGraphLaunch(G->ge, (hipStream_t)stream));
    return QCS_OK;
}
