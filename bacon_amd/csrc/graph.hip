// graph.hip — a batched solve captured once into a HIP graph and replayed.
//
// For callers that solve the same ensemble layout again and again (parameter
// sweeps, optimisation loops) the host work of a launch — argument checks,
// occupancy query, workspace allocation — happens once, at capture.  The
// graph holds the launch sequence a device entry point issues: the setup
// kernel (queue counter reset and descriptor copy; a kernel node, so the
// graph records no memset) and the persistent solve kernel.  For BDF the
// Newton (Broyden) iterations run inside that kernel, so one replay is the
// whole solve, iterations included.  The workspace is allocated at capture
// and owned by the graph; replays read the descriptor's device buffers as
// they are at replay time, so new inputs written into them are solved.
#include <cstring>

#include "bg_internal.h"

struct bg_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *ws = nullptr;
};

namespace bg {

template <class F>
static int capture(size_t ws_bytes, F &&launch, bg_graph **out) {
    if (!out) return BG_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    bg_graph *g = new bg_graph;
    hipStream_t s = nullptr;
    int rc = BG_OK;
    do {
        if (hipMalloc(&g->ws, ws_bytes) != hipSuccess) {
            rc = BG_ERR_HIP;
            break;
        }
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            rc = BG_ERR_HIP;
            break;
        }
        if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess) {
            rc = BG_ERR_HIP;
            break;
        }
        rc = launch(s, g->ws);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &graph);
        if (rc == BG_OK && e != hipSuccess) rc = BG_ERR_HIP;
        g->graph = graph;
        if (rc != BG_OK) break;
        if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
            rc = BG_ERR_HIP;
            break;
        }
    } while (0);
    if (s) hipStreamDestroy(s);
    if (rc != BG_OK) {
        bg_graph_destroy(g);
        return rc;
    }
    *out = g;
    return BG_OK;
}

}  // namespace bg

using namespace bg;

extern "C" int bg_graph_capture_rk(int method, int rhs, int semantics, const bg_ensemble_desc *d,
                                   const bg_launch *launch, bg_graph **out) {
    if (!d || (method != BG_RK45 && method != BG_RK23)) return BG_ERR_INVALID_ARGUMENT;
    if (d->n == 0) return BG_ERR_INVALID_ARGUMENT;
    bg_launch lc;
    memset(&lc, 0, sizeof lc);
    if (launch) lc = *launch;
    const bg_ensemble_desc dd = *d;
    return capture(
        kLaunchWorkspaceBytes,
        [&](hipStream_t s, void *ws) {
            return rk_launch_ensemble(method, rhs, semantics, dd, lc, s, ws);
        },
        out);
}

extern "C" int bg_graph_capture_bdf(int method, int rhs, int semantics, int lu_bits,
                                    const bg_ensemble_desc *d, bg_graph **out) {
    if (!d || (method != BG_BDF6 && method != BG_BDF2)) return BG_ERR_INVALID_ARGUMENT;
    if (lu_bits != 64 && lu_bits != 32) return BG_ERR_INVALID_ARGUMENT;
    if (d->n == 0) return BG_ERR_INVALID_ARGUMENT;
    const bg_ensemble_desc dd = *d;
    size_t bytes = 0;
    const int rc = bdf_workspace_bytes(method, rhs, semantics, lu_bits == 32, dd, &bytes);
    if (rc != BG_OK) return rc;
    return capture(
        bytes,
        [&](hipStream_t s, void *ws) {
            return bdf_launch_ensemble(method, rhs, semantics, lu_bits == 32, dd, s, ws, bytes);
        },
        out);
}

extern "C" int bg_graph_launch(bg_graph *g, void *stream) {
    if (!g || !g->exec) return BG_ERR_INVALID_ARGUMENT;
    BG_HIP_CHECK(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return BG_OK;
}

extern "C" void bg_graph_destroy(bg_graph *g) {
    if (!g) return;
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    if (g->ws) hipFree(g->ws);
    delete g;
}
