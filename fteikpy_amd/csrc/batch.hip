// fteik_solve_batch: many independent 2D and 3D models in one call.  The 3D models share
// every sweep launch (sweep3d.hip); the 2D models, one lane per source in the reference's
// serial order (sweep2d.hip), run meanwhile on a library-owned stream that is joined back
// into the caller's stream before the call returns.
#include "batch_internal.hpp"
#include "fteik_amd.h"
#include "fteik_device.hpp"

#include <mutex>
#include <vector>

namespace fteik {
namespace {

size_t align_up(size_t n) { return (n + 255) & ~(size_t)255; }

// One side stream per device, created on first use at the default priority and kept until
// the process ends.
hipStream_t side_stream(int *status) {
    static std::mutex mu;
    static std::vector<hipStream_t> streams;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *status = hip_status(e);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if ((int)streams.size() <= dev) streams.resize(dev + 1, nullptr);
    if (!streams[dev]) {
        e = hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking);
        if (e != hipSuccess) {
            streams[dev] = nullptr;
            *status = hip_status(e);
            return nullptr;
        }
    }
    *status = 0;
    return streams[dev];
}

}  // namespace
}  // namespace fteik

using namespace fteik;

extern "C" size_t fteik_batch_workspace_size(const fteik_problem *probs, int nprob) {
    if (!probs || nprob < 1) return 0;
    return align_up(batch3_bytes(probs, nprob)) + align_up(batch2_bytes(probs, nprob));
}

extern "C" int fteik_solve_batch(const fteik_problem *probs, int nprob, int nsweep, int max_sets,
                                 int want_grad, int group, void *workspace,
                                 size_t workspace_bytes, int use_graph, int *failed,
                                 void *stream_) {
    if (failed) *failed = -1;
    if (!probs || nprob < 1 || !workspace) return FTEIK_ERR_ARG;
    if (nsweep <= 0 && max_sets < 1) return FTEIK_ERR_ARG;
    int n2 = 0, n3 = 0;
    for (int i = 0; i < nprob; ++i) {
        if (probs[i].ndim == 2) ++n2;
        else if (probs[i].ndim == 3) ++n3;
        else return FTEIK_ERR_ARG;
    }
    if (workspace_bytes < fteik_batch_workspace_size(probs, nprob)) return FTEIK_ERR_WORKSPACE;
    char *ws3 = (char *)workspace;
    char *ws2 = ws3 + align_up(batch3_bytes(probs, nprob));
    hipStream_t stream = (hipStream_t)stream_;

    // every check before any device work
    Batch3 *b3 = nullptr;
    Batch2 *b2 = nullptr;
    int st = batch3_prepare(probs, nprob, nsweep, max_sets, want_grad, group, ws3, &b3, failed);
    if (!st) st = batch2_prepare(probs, nprob, nsweep, max_sets, want_grad, ws2, &b2, failed);
    if (st) {
        if (b3) batch3_free(b3);
        return st;
    }

    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    if (n2 && n3) {
        side = side_stream(&st);
        if (!st) st = hip_status(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        if (!st) st = hip_status(hipEventCreateWithFlags(&join, hipEventDisableTiming));
        if (!st) st = hip_status(hipEventRecord(fork, stream));
        if (!st) st = hip_status(hipStreamWaitEvent(side, fork, 0));
    }
    // 2D first (one long kernel on the side stream), then the 3D sets on the caller's
    // stream; the caller's stream waits for the 2D work at the end
    if (!st) st = batch2_enqueue(b2, side ? side : stream);
    if (!st && side) st = hip_status(hipEventRecord(join, side));
    if (!st) st = batch3_run(b3, use_graph, stream);
    if (side && join) {
        // join even after a failure, so nothing of this call is left running unordered
        const int sj = hip_status(hipStreamWaitEvent(stream, join, 0));
        if (!st) st = sj;
    }
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    batch3_free(b3);
    batch2_free(b2);
    return st;
}
