/*
 * fteik_amd: C ABI of the MI355X (gfx950) fast-sweeping eikonal library.
 *
 * Every entry point takes plain pointers and sizes, no C++ or torch types.  Device
 * pointers are HIP device memory (e.g. torch tensors' data_ptr()); "host" pointers are
 * ordinary host memory read before the call returns.  `stream` is a hipStream_t (NULL =
 * the default stream); every call is stream-ordered and only enqueues work unless stated.
 *
 * The reference (fteikpy 2.4.0) is pure Python/numba with no FFI; each entry point below
 * replaces one numba-compiled function called from the reference's Python layer, cited
 * as file:line of fteikpy/.  Array conventions are the reference's: slowness is cell based
 * (nz, nx[, ny]) C-ordered FP64, traveltime is node based (nz+1, nx+1[, ny+1]).
 *
 * Return codes: 0 OK; 1 source out of bound; 2 end point out of bound; 3 maximum number
 * of steps reached; 4 invalid argument; 5 workspace too small; 6 a dependency wait of the
 * persistent sweep gave up (result invalid); 7 unsupported option; negative: -(hipError_t).
 */
#ifndef FTEIK_AMD_H
#define FTEIK_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FTEIK_OK 0
#define FTEIK_ERR_SOURCE_OOB 1
#define FTEIK_ERR_ENDPOINT_OOB 2
#define FTEIK_ERR_MAX_STEP 3
#define FTEIK_ERR_ARG 4
#define FTEIK_ERR_WORKSPACE 5
#define FTEIK_ERR_TIMEOUT 6 /* persistent schedule: a grid-barrier wait gave up */
#define FTEIK_ERR_UNSUPPORTED 7
#define FTEIK_ERR_UNSETTLED 8

/* Library version string. */
const char *fteik_version(void);

/* ---------------------------------------------------------------------------------- */
/* 3D solve                                                                            */
/* ---------------------------------------------------------------------------------- */

/* Device workspace (bytes) fteik3d_solve needs for nsrc sources on an nz x nx x ny
 * (cells) grid: per-source bookkeeping plus one byte per node and source. */
size_t fteik3d_workspace_size(int nz, int nx, int ny, int nsrc);

/*
 * Replaces solve3d / fteik3d / fteik3d_vectorized (fteikpy/_fteik/_fteik3d.py:394-525),
 * called from Eikonal3D.solve (fteikpy/_solver.py:120-126).
 *
 * slow      device [nz, nx, ny] slowness (1/velocity), shared by all sources; one cell
 *           plane (nx * ny) must stay below 2^31 cells (else FTEIK_ERR_ARG)
 * dz,dx,dy  grid spacings, each in [1e-100, 1e100] (else FTEIK_ERR_ARG): the operator
 *           divides by fixed sums of inverse squared spacings through their reciprocals,
 *           which needs them to be normal doubles
 * src       host   [nsrc, 3] source coordinates relative to the grid origin
 * nsweep    > 0: exactly nsweep sweep sets (the reference's nsweep);
 *           <= 0: sets until one changes no node, at most max_sets
 * want_grad record the upwind direction of every update (needed by fteik3d_gradient)
 * tile_a    nodes per brick along z (0 = library default: from 32 sources 20..28 planes,
 *           the fewest steps down a column; 16 from 16 sources; else 8)
 * group     sources per workgroup (1, 2, 4, 8; 0 = default: 1 below 8 sources, else 2,
 *           except from 32 sources with schedule 1 or 2, where the first 2 sweep sets run
 *           4 per workgroup and the later ones 2; the group does not change the result)
 * schedule  2: overlapped passes: each launch runs every brick, of any pass, whose
 *           upwind bricks of its pass and whose tile's readers of the previous pass ran
 *           in earlier launches (508 launches per set instead of 688 for 160x320x320
 *           cells in 27-plane bricks); bricks under 8 planes run as schedule 1;
 *           1: one launch per (pass, brick anti-diagonal), ordered by the stream;
 *           0: persistent: one cooperative launch per sweep set (every workgroup
 *           resident) walking schedule 2's launch list (schedule 1's for bricks under 8
 *           planes) with a grid-wide barrier between entries; FTEIK_ERR_UNSUPPORTED on a
 *           device without cooperative launches, FTEIK_ERR_TIMEOUT (result invalid) if a
 *           barrier wait outlasted 5 s; use_graph is ignored; the call synchronizes the
 *           stream at the end (fixed nsweep) to read the barrier's give-up flag;
 *           0, 1 and 2 give identical results
 * tt        device [nsrc, nz+1, nx+1, ny+1] traveltimes (output)
 * sgn       device [nsrc, nz+1, nx+1, ny+1] uint8 packed upwind signs or NULL
 * vzero     device [nsrc] slowness of each source cell (output)
 * sets_done device [nsrc] sweep sets executed per source (output)
 * workspace device, at least fteik3d_workspace_size(nsrc) bytes
 * use_graph 1: capture a sweep set as a hipGraph and replay it (the default group from 32
 *           sources captures two, one per group); 0: launch directly;
 *           a captured set is kept (see fteik3d_graph_cache) and replayed by later calls
 *           whose launches take the same arguments: same device, slow, tt, sgn, sets_done
 *           and workspace pointers, grid, spacing, source count, brick height, group,
 *           gradient flag and schedule (nsweep and max_sets may differ)
 * Checks all sources before enqueuing anything: returns 1 ("source out of bound") with no
 * partial output, as the reference raises (fteikpy/_fteik/_fteik3d.py:402-407).
 * In converge mode (nsweep <= 0) the call synchronizes the stream once per set; the
 * persistent schedule synchronizes once at the end to read its error flag.
 */
int fteik3d_solve(const double *slow, int nz, int nx, int ny, double dz, double dx, double dy,
                  const double *src, int nsrc, int nsweep, int max_sets, int want_grad,
                  int tile_a, int group, int schedule, double *tt, uint8_t *sgn, double *vzero,
                  int *sets_done, void *workspace, size_t workspace_bytes, int use_graph,
                  void *stream);

/* No counterpart in the reference (instrumentation): launches of the sweep kernel per
 * sweep set that fteik3d_solve makes for this grid, source count, brick height (0 = the
 * default) and schedule; 0 for invalid arguments. */
int fteik3d_sweep_launches(int nz, int nx, int ny, int nsrc, int tile_a, int schedule);

/* No counterpart in the reference (resource control for callers that solve repeatedly):
 * the set graphs fteik3d_solve keeps for later calls (at most 8, least recently used
 * dropped).  clear != 0 destroys every kept graph (each after its last replay finished).
 * Returns the number of graphs kept; hits / misses (either may be NULL) receive the
 * process's counts of calls that found / did not find a kept graph. */
int fteik3d_graph_cache(int clear, long long *hits, long long *misses);

/* Bytes of the file-scope device scratch array of fteik3d_solve_noalloc (64 MiB). */
size_t fteik3d_scratch_capacity(void);

/*
 * No counterpart in the reference: fteik3d_solve for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The call allocates nothing: its workspace is a file-scope device array of
 * fteik3d_scratch_capacity() bytes that each device running the library holds until the
 * process ends (its address is looked up once per device on the first call there).
 * Arguments and results as fteik3d_solve with schedule 2 and use_graph 0; problems whose
 * fteik3d_workspace_size exceeds the capacity return FTEIK_ERR_WORKSPACE before anything
 * is enqueued.  All calls on one device share the array: calls that overlap on one device
 * (concurrent host threads, or different streams) race and must be serialized by the
 * caller.  Like fteik3d_solve, it reads `src` from host memory and synchronizes the stream
 * before its first kernel (and once per set when converging).
 */
int fteik3d_solve_noalloc(const double *slow, int nz, int nx, int ny, double dz, double dx,
                          double dy, const double *src, int nsrc, int nsweep, int max_sets,
                          int want_grad, int tile_a, int group, double *tt, uint8_t *sgn,
                          double *vzero, int *sets_done, void *stream);

/*
 * Replaces the gradient post-pass of fteik3d (fteikpy/_fteik/_fteik3d.py:464-488) that
 * fills TraveltimeGrid3D.gradient (fteikpy/_grid.py:267-278).
 * tt, sgn: device, one source ([nz+1, nx+1, ny+1]); vzero: that source's value;
 * src: host [3] relative to the origin.  ttgrad: device [nz+1, nx+1, ny+1, 3] (output).
 */
int fteik3d_gradient(const double *tt, const uint8_t *sgn, int nz, int nx, int ny, double dz,
                     double dx, double dy, const double *src, double vzero, double *ttgrad,
                     void *stream);

/* ---------------------------------------------------------------------------------- */
/* 3D adjoint-state gradient                                                           */
/* ---------------------------------------------------------------------------------- */

/* Bytes of a sweep-history buffer holding `events` node changes over at most max_passes
 * passes (8 per sweep set). */
size_t fteik3d_history_size(long long events, int max_passes);

/*
 * No counterpart in the reference: fteik3d_solve (exact schedule, no gradient signs) that
 * also records its sweep history -- every node change (node, time before the change),
 * pass by pass -- into `history` for fteik3d_adjoint_history.  Arguments as
 * fteik3d_solve; `history` is device memory of history_bytes (fteik3d_history_size with
 * max_passes = 8 * nsweep, or 8 * max_sets when converging).  *events receives the number
 * of changes; FTEIK_ERR_WORKSPACE means the buffer was too small for them (tt is still
 * correct; retry with fteik3d_history_size(*events, ...)).  nsrc * (nz+1)(nx+1)(ny+1) must
 * be below 2^32.
 */
int fteik3d_solve_logged(const double *slow, int nz, int nx, int ny, double dz, double dx,
                         double dy, const double *src, int nsrc, int nsweep, int max_sets,
                         int tile_a, int group, double *tt, double *vzero, int *sets_done,
                         void *workspace, size_t workspace_bytes, void *history,
                         size_t history_bytes, long long *events, int use_graph, void *stream);

/* Device workspace (bytes) of fteik3d_adjoint_history for a history of `events` changes. */
size_t fteik3d_adjoint_history_workspace_size(int nz, int nx, int ny, int nsrc,
                                              long long events);

/*
 * No counterpart in the reference: the exact gradient of a receiver-traveltime loss L with
 * respect to slowness -- reverse mode through the sweep history of fteik3d_solve_logged --
 * for T_rec[s, r] = vinterp3d of the solved grids (TraveltimeGrid3D.__call__,
 * fteikpy/_grid.py:193-219, on Eikonal3D.solve grids, fteikpy/_solver.py:120-126).
 * Passes are reversed last to first; the winning candidate of every logged change is
 * recomputed on the grid as it stood after its pass.
 *
 * tt       device [nsrc, nz+1, nx+1, ny+1]: the grids the logged solve returned; they are
 *          rolled back in place to the initial state (the caller keeps a copy if needed)
 * history  the logged solve's history buffer (read only)
 * origin, src, rec, dldt, dlds as fteik3d_adjoint (src absolute, host; dlds overwritten)
 * Returns 0, FTEIK_ERR_ARG (history of another grid / source count),
 * FTEIK_ERR_WORKSPACE (workspace too small or incomplete history), FTEIK_ERR_SOURCE_OOB,
 * FTEIK_ERR_UNSETTLED (a logged change is not reproduced by its candidate) or a negative
 * HIP error.  Synchronizes the stream once per pass that changed nodes.
 */
int fteik3d_adjoint_history(const double *slow, int nz, int nx, int ny, double dz, double dx,
                            double dy, const double *origin, const double *src, int nsrc,
                            double *tt, const void *history, const double *rec, int nrec,
                            const double *dldt, double *dlds, void *workspace,
                            size_t workspace_bytes, void *stream);

/* Device workspace (bytes) fteik3d_adjoint needs: 21 bytes per node and source. */
size_t fteik3d_adjoint_workspace_size(int nz, int nx, int ny, int nsrc);

/*
 * No counterpart in the reference: the gradient of a receiver-traveltime loss L with
 * respect to slowness, for T_rec[s, r] = vinterp3d of converged fteik3d grids (what
 * TraveltimeGrid3D.__call__ returns, fteikpy/_grid.py:193-219, on the grids of
 * Eikonal3D.solve, fteikpy/_solver.py:120-126).  The upwind stencil of every node is
 * recomputed from the converged times and the adjoint system is solved exactly in reverse
 * dependency order.
 *
 * slow    device [nz, nx, ny] slowness the grids were solved with
 * origin  host [3] grid origin; src: host [nsrc, 3] ABSOLUTE source coordinates
 * tt      device [nsrc, nz+1, nx+1, ny+1] converged traveltimes (fteik3d_solve, nsweep <= 0)
 * rec     device [nrec, 3] absolute receiver coordinates (shared by all sources)
 * dldt    device [nsrc, nrec] dL/dT_rec
 * dlds    device [nz, nx, ny] output dL/dslow (overwritten)
 * stale   host out: nodes carrying adjoint whose held time no current candidate reproduces
 *         (0 when the sweeps reached a consistent fixed point); may be NULL
 * rounds  host out: dependency rounds executed; may be NULL
 * Returns 0, FTEIK_ERR_SOURCE_OOB, FTEIK_ERR_ARG (also for nsrc * nodes >= 2^32),
 * FTEIK_ERR_WORKSPACE, FTEIK_ERR_UNSETTLED (the stencil graph has a cycle) or a negative
 * HIP error.  Synchronizes the stream once per dependency round.
 */
int fteik3d_adjoint(const double *slow, int nz, int nx, int ny, double dz, double dx, double dy,
                    const double *origin, const double *src, int nsrc, const double *tt,
                    const double *rec, int nrec, const double *dldt, double *dlds, int *stale,
                    int *rounds, void *workspace, size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------------------- */
/* 2D solve                                                                            */
/* ---------------------------------------------------------------------------------- */

size_t fteik2d_workspace_size(int nsrc);

/*
 * Replaces solve2d / fteik2d / fteik2d_vectorized (fteikpy/_fteik/_fteik2d.py:290-674),
 * called from Eikonal2D.solve (fteikpy/_solver.py:48-54).  Same conventions as
 * fteik3d_solve with src [nsrc, 2] and tt [nsrc, nz+1, nx+1].  The reference's 2D set is
 * a single serial dependency chain (columns interleave the +z and -z passes), so each
 * source gets the results of the reference's exact order; sources run in parallel, one
 * wave each, which stages the two active columns ((nz+2) x 72 B) and evaluates each z pass
 * in speculative blocks of 64 nodes that finalize only provably exact nodes.  The staged
 * rows sit in LDS when they fit one workgroup's (nz <= 2273 on MI355X); deeper models get
 * them in global memory the call allocates stream-ordered (hipMallocAsync) and frees.  dz, dx in [1e-100, 1e100]
 * (else FTEIK_ERR_ARG), as for fteik3d_solve.
 */
int fteik2d_solve(const double *slow, int nz, int nx, double dz, double dx, const double *src,
                  int nsrc, int nsweep, int max_sets, int want_grad, double *tt, uint8_t *sgn,
                  double *vzero, int *sets_done, void *workspace, size_t workspace_bytes,
                  void *stream);

/* Replaces the 2D gradient post-pass (fteikpy/_fteik/_fteik2d.py:623-639). */
int fteik2d_gradient(const double *tt, const uint8_t *sgn, int nz, int nx, double dz, double dx,
                     const double *src, double vzero, double *ttgrad, void *stream);

/* ---------------------------------------------------------------------------------- */
/* Batch of independent models                                                          */
/* ---------------------------------------------------------------------------------- */

/* One model of a batch (fteik_solve_batch): the arguments of one fteik3d_solve or
 * fteik2d_solve call.  ndim 2: ny and dy are ignored, src is [nsrc, 2] and the grids are
 * [nsrc, nz+1, nx+1]. */
typedef struct fteik_problem {
    int ndim;             /* 2 or 3 */
    int nz, nx, ny;       /* cells */
    double dz, dx, dy;
    const double *slow;   /* device [nz, nx(, ny)] slowness */
    const double *src;    /* host [nsrc, ndim] sources relative to the model's origin */
    int nsrc;
    int tile_a;           /* 3D brick height along z (0 = the library default for the
                             batch's total number of 3D sources, see fteik3d_solve) */
    double *tt;           /* device [nsrc, nz+1, nx+1(, ny+1)] traveltimes (output) */
    uint8_t *sgn;         /* device upwind signs, shaped as tt (uint8); NULL without want_grad */
    double *vzero;        /* device [nsrc] (output) */
    int *sets_done;       /* device [nsrc] (output) */
} fteik_problem;

/* Device workspace (bytes) of fteik_solve_batch for these problems. */
size_t fteik_batch_workspace_size(const fteik_problem *probs, int nprob);

/*
 * No single counterpart in the reference, which solves one model per Eikonal2D/3D.solve
 * call (fteikpy/_solver.py:48-54, :120-126; solve2d / solve3d, _fteik2d.py:667-674,
 * _fteik3d.py:516-525): solves many independent models of uneven size in one call.  Each
 * problem's outputs equal those of its own fteik3d_solve / fteik2d_solve call bit for bit.
 * The 3D models share every sweep launch -- one launch per (pass, brick anti-diagonal)
 * covers the bricks of all models and their active sources -- and converge per source;
 * the 2D models (one wave per source, the reference's results per z pass) run
 * concurrently on a library-owned stream joined back into `stream` (they overlap the 3D
 * sweeps only while that stream has a hardware queue of its own: HIP maps a process's
 * streams onto GPU_MAX_HW_QUEUES queues, 4 by default).
 *
 * nsweep, max_sets  as fteik3d_solve (nsweep <= 0: converge, at most max_sets sets)
 * want_grad         record upwind signs: every problem must then pass `sgn`
 * group             3D sources per workgroup (0 = default, 1 in a batch; 1, 2, 4, 8), the
 *                   same for all
 * failed            host out (may be NULL): index of the first problem with a source out
 *                   of bound, else -1
 * Checks every source of every problem before enqueuing anything (returns 1, "source out
 * of bound").  In converge mode the call synchronizes `stream` once per set.
 */
int fteik_solve_batch(const fteik_problem *probs, int nprob, int nsweep, int max_sets,
                      int want_grad, int group, void *workspace, size_t workspace_bytes,
                      int use_graph, int *failed, void *stream);

/* ---------------------------------------------------------------------------------- */
/* Grid preparation (smoothing, resampling)                                            */
/* ---------------------------------------------------------------------------------- */

/*
 * One axis of BaseGrid2D/3D.smooth (fteikpy/_base.py:114-123, :203-212), which calls
 * scipy.ndimage.gaussian_filter: correlation of `in` (device, C-ordered `shape`, ndim 1..3)
 * along `axis` with the symmetric kernel `weights` (device, 2 radius + 1 values) and
 * 'reflect' boundaries, into `out` (device, same shape, not aliasing `in`).  A Gaussian
 * filter is one call per axis, axis 0 first, each reading the previous output.
 */
int fteik_correlate1d(const double *in, double *out, int ndim, const int *shape, int axis,
                      const double *weights, int radius, void *stream);

/* Device workspace (bytes) of fteik_resample. */
size_t fteik_resample_workspace_size(int ndim, const int *new_shape);

/*
 * BaseGrid2D/3D.resample (fteikpy/_base.py:81-112, :165-201), which evaluates
 * scipy.interpolate.RegularGridInterpolator (method 0 'linear', 1 'nearest') of the grid
 * `in` (device, `shape` nodes at origin + gridsize * i) on the nodes of
 * linspace(axis[0], axis[-1], new_shape[a]) per axis, into `out` (device, new_shape).
 * origin, gridsize, shape, new_shape: host [ndim], ndim 2 or 3, every shape[a] >= 2.
 * Synchronizes the stream once (upload of the per-axis cell indices).
 */
int fteik_resample(const double *in, int ndim, const int *shape, const double *origin,
                   const double *gridsize, const int *new_shape, int method, double *out,
                   void *workspace, size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------------------- */
/* Interpolation                                                                       */
/* ---------------------------------------------------------------------------------- */

/*
 * Trilinear interpolation, replaces interp3d (fteikpy/_interp/_interp3d.py:7-195), called
 * from BaseGrid3D.__call__ (fteikpy/_base.py:139-163).
 * v: device grid [n0, n1, n2] read with element stride `stride` (1 for a plain grid, 3 for
 * one component of an interleaved gradient grid).  origin, d: host [3].  q: device [nq, 3].
 * out: device [nq]; fill for points outside the grid.
 */
int fteik3d_interp(const double *v, int n0, int n1, int n2, int stride, const double *origin,
                   const double *d, const double *q, int nq, double fill, double *out,
                   void *stream);

/*
 * Apparent-velocity trilinear traveltime interpolation, replaces vinterp3d
 * (fteikpy/_interp/_vinterp3d.py:7-291), called from TraveltimeGrid3D.__call__
 * (fteikpy/_grid.py:193-219).  tt: device [n0, n1, n2]; src: host [3] ABSOLUTE source
 * coordinates; vzero: slowness of the source cell.
 */
int fteik3d_vinterp(const double *tt, int n0, int n1, int n2, const double *origin,
                    const double *d, const double *q, int nq, const double *src, double vzero,
                    double fill, double *out, void *stream);

/* 2D twins: interp2d (fteikpy/_interp/_interp2d.py:7-95), vinterp2d
 * (fteikpy/_interp/_vinterp2d.py:7-131).  origin, d, src: host [2]; q: device [nq, 2]. */
int fteik2d_interp(const double *v, int n0, int n1, int stride, const double *origin,
                   const double *d, const double *q, int nq, double fill, double *out,
                   void *stream);
int fteik2d_vinterp(const double *tt, int n0, int n1, const double *origin, const double *d,
                    const double *q, int nq, const double *src, double vzero, double fill,
                    double *out, void *stream);

/* ---------------------------------------------------------------------------------- */
/* Ray tracing                                                                         */
/* ---------------------------------------------------------------------------------- */

/*
 * A-posteriori ray tracing, replaces ray3d (fteikpy/_fteik/_ray3d.py:9-208), called from
 * TraveltimeGrid3D.raytrace (fteikpy/_grid.py:221-265).
 * grad: device [n0, n1, n2, 3] normalized gradient (fteik3d_gradient output).
 * p: device [n, 3] end points (receivers), src: host [3] ABSOLUTE source.
 * rays: device [n, max_step, 3] work buffer in receiver-to-source order (as the
 * reference's buffer); counts: device [n] index of the last point; status: device [n]
 * per-ray code (0, 2 end point out of bound, 3 maximum number of steps reached).
 * Enqueues only: the caller reads status after synchronizing.
 */
int fteik3d_raytrace(const double *grad, int n0, int n1, int n2, const double *origin,
                     const double *d, const double *p, int n, const double *src,
                     double stepsize, int max_step, int honor_grid, double *rays, int *counts,
                     int *status, void *stream);

/*
 * No counterpart in the reference, which traces the rays of one traveltime grid per call
 * (TraveltimeGrid3D.raytrace, fteikpy/_grid.py:221-265): the rays of several traveltime
 * grids of one model (same n0, n1, n2, origin and d; one grid per source) in one launch,
 * each ray exactly as fteik3d_raytrace traces it.  grads: device [ngrid] pointers to the
 * grids' gradient arrays ([n0, n1, n2, 3] each); grid_of_ray: device [n], the grid each
 * ray follows (a value outside [0, ngrid) gives that ray status 4, the argument error);
 * src: device [ngrid, 3] ABSOLUTE sources; p, rays, counts, status as fteik3d_raytrace.
 * For callers with many sources of one model (one wave per 64 rays instead of a
 * latency-bound launch per source).
 */
int fteik3d_raytrace_batch(const double *const *grads, const int *grid_of_ray,
                           const double *src, int ngrid, int n0, int n1, int n2,
                           const double *origin, const double *d, const double *p, int n,
                           double stepsize, int max_step, int honor_grid, double *rays,
                           int *counts, int *status, void *stream);

/* 2D twin of fteik3d_raytrace, replaces ray2d (fteikpy/_fteik/_ray2d.py:9-153). */
int fteik2d_raytrace(const double *grad, int n0, int n1, const double *origin, const double *d,
                     const double *p, int n, const double *src, double stepsize, int max_step,
                     int honor_grid, double *rays, int *counts, int *status, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FTEIK_AMD_H */
