// A-posteriori ray tracing on MI355X: one lane per ray, stepping from the receiver down the
// interpolated unit traveltime gradient until the source is within one step
// (fteikpy/_fteik/_ray3d.py:9-120, _ray2d.py:9-91, shrink: _fteik/_common.py:4-28).
#include "fteik_amd.h"
#include "interp_device.hpp"

namespace fteik {
namespace {

// Step fraction that stops the ray on the first cell face it would cross
// (fteikpy/_fteik/_common.py:4-28).
template <int ND>
__device__ inline double shrink(const double *pcur, const double *delta, const double *lower,
                                const double *upper) {
    bool condl = false, condu = false;
    double bl = 0.0, bu = 0.0;
    for (int a = 0; a < ND; ++a) {
        const double tmp = pcur[a] - delta[a];
        if (tmp < lower[a]) {
            const double b = (pcur[a] - lower[a]) / delta[a];
            if (!condl || b < bl) bl = b;
            condl = true;
        }
        if (tmp > upper[a]) {
            const double b = (pcur[a] - upper[a]) / delta[a];
            if (!condu || b < bu) bu = b;
            condu = true;
        }
    }
    if (condl && condu) return pymin(bl, bu);
    if (condl) return bl;
    if (condu) return bu;
    return 1.0;
}

// Lower/upper cell bounds around p on one axis; a point on a node takes the cell below
// it (_ray3d.py:39-47, :90-98).  The upper index is clamped to the last node where the
// reference would index past the axis.
__device__ __forceinline__ void bounds(const Axis &ax, double p, double *lo, double *hi, int *idx) {
    const int i = ax.cell(p);
    *idx = i;
    *lo = (p == ax.at(i)) ? ax.at(i > 0 ? i - 1 : 0) : ax.at(i);
    *hi = ax.at(i + 1 < ax.n ? i + 1 : ax.n - 1);
}

// One ray from the receiver pcur down the gradient grid `grad` to the source srcp: writes
// its points to ray[0 ..] (receiver first, the source last) and their number minus one to
// *count; returns 0 or the ray's error code.
template <int ND>
__device__ int trace_ray(const double *grad, const Axis (&ax)[3], const double (&srcp)[3],
                         double (&pcur)[3], double stepsize, int max_step, int honor_grid,
                         double *ray, int *count_out) {
    double delta[3], lower[3], upper[3];
    *count_out = 0;
    for (int a = 0; a < ND; ++a)
        if (!ax[a].contains(pcur[a])) return FTEIK_ERR_ENDPOINT_OOB;
    int isrc[3] = {0, 0, 0};
    if (honor_grid) {
        int dummy;
        for (int a = 0; a < ND; ++a) {
            bounds(ax[a], pcur[a], &lower[a], &upper[a], &dummy);
            isrc[a] = ax[a].cell(srcp[a]);
        }
    }
    auto dist = [&]() {
        return ND == 3 ? dist3(srcp[0], srcp[1], srcp[2], pcur[0], pcur[1], pcur[2])
                       : dist2(srcp[0], srcp[1], pcur[0], pcur[1]);
    };
    int count = 1;
    for (int a = 0; a < ND; ++a) ray[a] = pcur[a];
    while (dist() >= stepsize) {
        double g[3];
        for (int a = 0; a < ND; ++a)
            g[a] = ND == 3 ? interp3(ax[0], ax[1], ax[2], grad + a, 3, pcur[0], pcur[1], pcur[2], NAN)
                           : interp2(ax[0], ax[1], grad + a, 2, pcur[0], pcur[1], NAN);
        const double gn = ND == 3 ? sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2])
                                  : sqrt(g[0] * g[0] + g[1] * g[1]);
        double gni;
        if (gn > 0.0) gni = 1.0 / gn;
        else break;
        for (int a = 0; a < ND; ++a) delta[a] = stepsize * g[a] * gni;
        if (honor_grid) {
            const double fac = shrink<ND>(pcur, delta, lower, upper);
            for (int a = 0; a < ND; ++a) {
                pcur[a] -= fac * delta[a];
                pcur[a] = pymin(pymax(pcur[a], ax[a].at(0)), ax[a].last());
            }
            if (fac < 1.0) {
                // grid magnetism, _ray3d.py:81-88
                for (int a = 0; a < ND; ++a) {
                    if (fabs(pcur[a] - lower[a]) < 1.0e-8) pcur[a] = lower[a];
                    else if (fabs(pcur[a] - upper[a]) < 1.0e-8) pcur[a] = upper[a];
                }
                int idx[3] = {0, 0, 0};
                for (int a = 0; a < ND; ++a) bounds(ax[a], pcur[a], &lower[a], &upper[a], &idx[a]);
                for (int a = 0; a < ND; ++a) ray[ND * count + a] = pcur[a];
                ++count;
                bool at_src = true;
                for (int a = 0; a < ND; ++a) at_src = at_src && idx[a] == isrc[a];
                if (at_src) break;
            }
        } else {
            for (int a = 0; a < ND; ++a) {
                pcur[a] -= delta[a];
                pcur[a] = pymin(pymax(pcur[a], ax[a].at(0)), ax[a].last());
            }
            for (int a = 0; a < ND; ++a) ray[ND * count + a] = pcur[a];
            ++count;
        }
        if (count >= max_step) return FTEIK_ERR_MAX_STEP;
    }
    for (int a = 0; a < ND; ++a) ray[ND * count + a] = srcp[a];
    *count_out = count;
    return 0;
}

template <int ND>
__global__ void ray_kernel(const double *grad, Axis A0, Axis A1, Axis A2, const double *p, int n,
                           double s0, double s1, double s2, double stepsize, int max_step,
                           int honor_grid, double *rays, int *counts, int *status) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const Axis ax[3] = {A0, A1, A2};
    const double srcp[3] = {s0, s1, s2};
    double pcur[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < ND; ++a) pcur[a] = p[ND * r + a];
    status[r] = trace_ray<ND>(grad, ax, srcp, pcur, stepsize, max_step, honor_grid,
                              rays + (long long)r * max_step * ND, counts + r);
}

// Rays of several traveltime grids of one model: ray r follows grid grid_of_ray[r], whose
// gradient array is grads[k] and source src[3 k ..].
__global__ void ray_batch_kernel(const double *const *grads, const int *grid_of_ray,
                                 const double *src, int ngrid, Axis A0, Axis A1, Axis A2,
                                 const double *p, int n, double stepsize, int max_step,
                                 int honor_grid, double *rays, int *counts, int *status) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int k = grid_of_ray[r];
    if (k < 0 || k >= ngrid) {  // not a grid of the batch
        counts[r] = 0;
        status[r] = FTEIK_ERR_ARG;
        return;
    }
    const Axis ax[3] = {A0, A1, A2};
    const double srcp[3] = {src[3 * k], src[3 * k + 1], src[3 * k + 2]};
    double pcur[3] = {p[3 * r], p[3 * r + 1], p[3 * r + 2]};
    status[r] = trace_ray<3>(grads[k], ax, srcp, pcur, stepsize, max_step, honor_grid,
                             rays + (long long)r * max_step * 3, counts + r);
}

}  // namespace
}  // namespace fteik

using namespace fteik;

extern "C" int fteik3d_raytrace(const double *grad, int n0, int n1, int n2, const double *origin,
                                const double *d, const double *p, int n, const double *src,
                                double stepsize, int max_step, int honor_grid, double *rays,
                                int *counts, int *status, void *stream) {
    if (!grad || !origin || !d || !p || !src || !rays || !counts || !status || n < 0 ||
        max_step < 2 || !(stepsize > 0.0) || n0 < 2 || n1 < 2 || n2 < 2)
        return FTEIK_ERR_ARG;
    if (n == 0) return 0;
    Axis A0{origin[0], d[0], n0}, A1{origin[1], d[1], n1}, A2{origin[2], d[2], n2};
    hipLaunchKernelGGL(ray_kernel<3>, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)stream, grad,
                       A0, A1, A2, p, n, src[0], src[1], src[2], stepsize, max_step, honor_grid,
                       rays, counts, status);
    return hip_status(hipGetLastError());
}

extern "C" int fteik3d_raytrace_batch(const double *const *grads, const int *grid_of_ray,
                                      const double *src, int ngrid, int n0, int n1, int n2,
                                      const double *origin, const double *d, const double *p,
                                      int n, double stepsize, int max_step, int honor_grid,
                                      double *rays, int *counts, int *status, void *stream) {
    if (!grads || !grid_of_ray || !src || ngrid < 1 || !origin || !d || !p || !rays || !counts ||
        !status || n < 0 || max_step < 2 || !(stepsize > 0.0) || n0 < 2 || n1 < 2 || n2 < 2)
        return FTEIK_ERR_ARG;
    if (n == 0) return 0;
    Axis A0{origin[0], d[0], n0}, A1{origin[1], d[1], n1}, A2{origin[2], d[2], n2};
    hipLaunchKernelGGL(ray_batch_kernel, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)stream,
                       grads, grid_of_ray, src, ngrid, A0, A1, A2, p, n, stepsize, max_step,
                       honor_grid, rays, counts, status);
    return hip_status(hipGetLastError());
}

extern "C" int fteik2d_raytrace(const double *grad, int n0, int n1, const double *origin,
                                const double *d, const double *p, int n, const double *src,
                                double stepsize, int max_step, int honor_grid, double *rays,
                                int *counts, int *status, void *stream) {
    if (!grad || !origin || !d || !p || !src || !rays || !counts || !status || n < 0 ||
        max_step < 2 || !(stepsize > 0.0) || n0 < 2 || n1 < 2)
        return FTEIK_ERR_ARG;
    if (n == 0) return 0;
    Axis A0{origin[0], d[0], n0}, A1{origin[1], d[1], n1}, A2{0.0, 1.0, 2};
    hipLaunchKernelGGL(ray_kernel<2>, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)stream, grad,
                       A0, A1, A2, p, n, src[0], src[1], 0.0, stepsize, max_step, honor_grid, rays,
                       counts, status);
    return hip_status(hipGetLastError());
}
