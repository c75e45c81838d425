// Internal interface between batch.hip (fteik_solve_batch) and the 3D / 2D solvers that
// plan and run their parts of a batch of independent models.
#pragma once

#include <hip/hip_runtime.h>

#include "fteik_amd.h"

namespace fteik {

struct Batch3;  // 3D models: sweep3d.hip
struct Batch2;  // 2D models: sweep2d.hip

// Workspace bytes of each part (0 when the batch holds no model of that dimension).
size_t batch3_bytes(const fteik_problem *probs, int nprob);
size_t batch2_bytes(const fteik_problem *probs, int nprob);

// Host-only argument and source checks and the launch plan; nothing is enqueued.  On a
// source out of bound, *failed receives the problem index.  *out is set on success only.
int batch3_prepare(const fteik_problem *probs, int nprob, int nsweep, int max_sets,
                   int want_grad, int group, char *ws, Batch3 **out, int *failed);
int batch2_prepare(const fteik_problem *probs, int nprob, int nsweep, int max_sets,
                   int want_grad, char *ws, Batch2 **out, int *failed);

// Enqueue (and, converging, drive) the part on `stream`.
int batch3_run(Batch3 *b, int use_graph, hipStream_t stream);
int batch2_enqueue(Batch2 *b, hipStream_t stream);

void batch3_free(Batch3 *b);
void batch2_free(Batch2 *b);

}  // namespace fteik
