// plan.hip -- products whose operands keep their structure and change values:
// the numeric step re-run as one captured hipGraph.  No reference counterpart
// (the reference recomputes everything per run, src/spgemm_raw_v2.cu:111-240).
//
// tsg_plan_create runs the product (or the batch) once and keeps what its
// numeric step needs: the row bins and links of the binned pipeline, and for
// a batch the block-diagonal operands and product with the job tables that copy
// values into and out of them.  tsg_plan_execute launches only kernels -- the
// value copies into the batch operands, the numeric kernels of every bin, the
// copies out to each C_i -- so the sequence is captured once into a hipGraph on
// the stream and replayed on later calls: one launch per execution, no host
// sync, no allocation.
#include <vector>

#include "rowstaged.h"

struct tsg_plan {
    tsg::Context *ctx = nullptr;
    int32_t count = 0;
    tsg_tile_matrix A{}, B{}, C{};  // single product: the operands and output (caller-owned arrays)
    tsg::NumericReplay nr;          // single product
    tsg::BatchKeep bk;              // batch
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace tsg {

// The kernels of one execution, launched on the context's stream.
static int plan_launch(tsg_plan *p) {
    Context *ctx = p->ctx;
    if (p->count == 1) return replay_numeric(ctx, &p->A, &p->B, &p->C, p->nr);
    BatchKeep &k = p->bk;
    if (k.nvin) TSG_CHECK(launch_jobs(ctx, k.vin, k.nvin, k.vin_max));
    TSG_CHECK(replay_numeric(ctx, &k.BA, &k.BB, &k.BC, k.nr));
    if (k.nvout) TSG_CHECK(launch_jobs(ctx, k.vout, k.nvout, k.vout_max));
    return TSG_OK;
}

}  // namespace tsg

using namespace tsg;

extern "C" {

int tsg_plan_create(tsg_context *c, int32_t count, const tsg_tile_matrix *A, const tsg_tile_matrix *B,
                    tsg_tile_matrix *C, const tsg_options *opts, tsg_plan **out) {
    Context *ctx = reinterpret_cast<Context *>(c);
    if (!ctx || count < 1 || !A || !B || !C || !out) return fail(TSG_ERR_ARG, "bad plan arguments");
    *out = nullptr;
    (void)hipSetDevice(ctx->device);
    tsg_plan *p = new tsg_plan();
    p->ctx = ctx;
    p->count = count;
    int rc;
    if (count == 1) {
        rc = check_operands(A, B);
        if (!rc) rc = spgemm_dispatch(ctx, A, B, C, opts, nullptr, &p->nr);
        if (!rc) {
            p->A = *A;
            p->B = *B;
            p->C = *C;
        }
    } else {
        rc = spgemm_batched(ctx, count, A, B, C, opts, nullptr, &p->bk);
    }
    if (rc) {
        free_plan(ctx, p->nr.p);
        delete p;
        return rc;
    }
    *out = p;
    return TSG_OK;
}

int tsg_plan_execute(tsg_plan *p) {
    if (!p || !p->ctx) return fail(TSG_ERR_ARG, "null plan");
    Context *ctx = p->ctx;
    (void)hipSetDevice(ctx->device);
    if (!p->exec) {  // first execution: capture the launches, then instantiate
        TSG_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int rc = plan_launch(p);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(TSG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
        p->graph = g;
        if (hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess) {
            p->exec = nullptr;
            return fail(TSG_ERR_HIP, "graph instantiation failed");
        }
    }
    TSG_HIP(hipGraphLaunch(p->exec, ctx->stream));
    return TSG_OK;
}

int tsg_plan_destroy(tsg_plan *p) {
    if (!p) return TSG_OK;
    Context *ctx = p->ctx;
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
    }
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (ctx) {
        free_plan(ctx, p->nr.p);
        free_batch_keep(ctx, p->bk);
    }
    delete p;
    return TSG_OK;
}

}  // extern "C"
