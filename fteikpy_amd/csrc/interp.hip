// Receiver traveltimes (apparent-velocity interpolation) and plain bi/trilinear
// interpolation on MI355X: one lane per query point.
//   vinterp3d  fteikpy/_interp/_vinterp3d.py:7-291     vinterp2d  _vinterp2d.py:7-131
//   interp3d   fteikpy/_interp/_interp3d.py:7-195      interp2d   _interp2d.py:7-95
#include "fteik_amd.h"
#include "interp_device.hpp"

namespace fteik {
namespace {

__global__ void interp3d_kernel(const double *v, Axis X, Axis Y, Axis Z, int stride,
                                const double *q, int nq, double fill, double *out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nq) return;
    out[n] = interp3(X, Y, Z, v, stride, q[3 * n], q[3 * n + 1], q[3 * n + 2], fill);
}

__global__ void interp2d_kernel(const double *v, Axis X, Axis Y, int stride, const double *q,
                                int nq, double fill, double *out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nq) return;
    out[n] = interp2(X, Y, v, stride, q[2 * n], q[2 * n + 1], fill);
}

// _vinterp3d.py:7-264.  Corner weights are d/t (distance from the source over traveltime,
// i.e. apparent slowness); the blend is inverted back into a time.
__global__ void vinterp3d_kernel(const double *tt, Axis X, Axis Y, Axis Z, const double *q,
                                 int nq, double xs, double ys, double zs, double vzero,
                                 double fill, double *out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nq) return;
    const double xq = q[3 * n], yq = q[3 * n + 1], zq = q[3 * n + 2];
    if (!(X.contains(xq) && Y.contains(yq) && Z.contains(zq))) {
        out[n] = fill;
        return;
    }
    const int i = X.cell(xq), j = Y.cell(yq), k = Z.cell(zq);
    if (X.cell(xs) == i && Y.cell(ys) == j && Z.cell(zs) == k) {
        out[n] = vzero * dist3(xs, ys, zs, xq, yq, zq);
        return;
    }
    const Corner cx = corner(X, i), cy = corner(Y, j), cz = corner(Z, k);
    const double x1 = cx.x1, x2 = cx.x2, y1 = cy.x1, y2 = cy.x2, z1 = cz.x1, z2 = cz.x2;
    auto V = [&](int a, int b, int c) { return tt[((long long)a * Y.n + b) * Z.n + c]; };
    double v111 = V(i, j, k), d111 = dist3(xs, ys, zs, x1, y1, z1);
    double v211 = 1.0, d211 = 0.0, v121 = 1.0, d121 = 0.0, v221 = 1.0, d221 = 0.0;
    double v112 = 1.0, d112 = 0.0, v212 = 1.0, d212 = 0.0, v122 = 1.0, d122 = 0.0;
    double v222 = 1.0, d222 = 0.0;
    if (cx.has2) { v211 = V(i + 1, j, k); d211 = dist3(xs, ys, zs, x2, y1, z1); }
    if (cy.has2) { v121 = V(i, j + 1, k); d121 = dist3(xs, ys, zs, x1, y2, z1); }
    if (cx.has2 && cy.has2) { v221 = V(i + 1, j + 1, k); d221 = dist3(xs, ys, zs, x2, y2, z1); }
    if (cz.has2) { v112 = V(i, j, k + 1); d112 = dist3(xs, ys, zs, x1, y1, z2); }
    if (cx.has2 && cz.has2) { v212 = V(i + 1, j, k + 1); d212 = dist3(xs, ys, zs, x2, y1, z2); }
    if (cy.has2 && cz.has2) { v122 = V(i, j + 1, k + 1); d122 = dist3(xs, ys, zs, x1, y2, z2); }
    if (cx.has2 && cy.has2 && cz.has2) {
        v222 = V(i + 1, j + 1, k + 1);
        d222 = dist3(xs, ys, zs, x2, y2, z2);
    }
    // a zero corner time means the source sits on that node (_vinterp3d.py:246-250)
    if (!(v111 != 0.0 && v211 != 0.0 && v121 != 0.0 && v221 != 0.0 && v112 != 0.0 &&
          v212 != 0.0 && v122 != 0.0 && v222 != 0.0)) {
        out[n] = vzero * dist3(xs, ys, zs, xq, yq, zq);
        return;
    }
    double vq = d111 / v111 * fabs((x2 - xq) * (y2 - yq) * (z2 - zq));
    vq += d211 / v211 * fabs((x1 - xq) * (y2 - yq) * (z2 - zq));
    vq += d121 / v121 * fabs((x2 - xq) * (y1 - yq) * (z2 - zq));
    vq += d221 / v221 * fabs((x1 - xq) * (y1 - yq) * (z2 - zq));
    vq += d112 / v112 * fabs((x2 - xq) * (y2 - yq) * (z1 - zq));
    vq += d212 / v212 * fabs((x1 - xq) * (y2 - yq) * (z1 - zq));
    vq += d122 / v122 * fabs((x2 - xq) * (y1 - yq) * (z1 - zq));
    vq += d222 / v222 * fabs((x1 - xq) * (y1 - yq) * (z1 - zq));
    vq /= fabs((x2 - x1) * (y2 - y1) * (z2 - z1));
    out[n] = dist3(xs, ys, zs, xq, yq, zq) / vq;
}

// _vinterp2d.py:7-108
__global__ void vinterp2d_kernel(const double *tt, Axis X, Axis Y, const double *q, int nq,
                                 double xs, double ys, double vzero, double fill, double *out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nq) return;
    const double xq = q[2 * n], yq = q[2 * n + 1];
    if (!(X.contains(xq) && Y.contains(yq))) {
        out[n] = fill;
        return;
    }
    const int i = X.cell(xq), j = Y.cell(yq);
    if (X.cell(xs) == i && Y.cell(ys) == j) {
        out[n] = vzero * dist2(xs, ys, xq, yq);
        return;
    }
    const Corner cx = corner(X, i), cy = corner(Y, j);
    const double x1 = cx.x1, x2 = cx.x2, y1 = cy.x1, y2 = cy.x2;
    auto V = [&](int a, int b) { return tt[(long long)a * Y.n + b]; };
    double v11 = V(i, j), d11 = dist2(xs, ys, x1, y1);
    double v21 = 1.0, d21 = 0.0, v12 = 1.0, d12 = 0.0, v22 = 1.0, d22 = 0.0;
    if (cx.has2) { v21 = V(i + 1, j); d21 = dist2(xs, ys, x2, y1); }
    if (cy.has2) { v12 = V(i, j + 1); d12 = dist2(xs, ys, x1, y2); }
    if (cx.has2 && cy.has2) { v22 = V(i + 1, j + 1); d22 = dist2(xs, ys, x2, y2); }
    if (!(v11 != 0.0 && v21 != 0.0 && v12 != 0.0 && v22 != 0.0)) {
        out[n] = vzero * dist2(xs, ys, xq, yq);
        return;
    }
    double vq = d11 / v11 * fabs((x2 - xq) * (y2 - yq));
    vq += d21 / v21 * fabs((x1 - xq) * (y2 - yq));
    vq += d12 / v12 * fabs((x2 - xq) * (y1 - yq));
    vq += d22 / v22 * fabs((x1 - xq) * (y1 - yq));
    vq /= fabs((x2 - x1) * (y2 - y1));
    out[n] = dist2(xs, ys, xq, yq) / vq;
}

bool axes_ok(const int *n, const double *d, int nd) {
    for (int a = 0; a < nd; ++a)
        if (n[a] < 2 || !(d[a] > 0.0)) return false;
    return true;
}

}  // namespace
}  // namespace fteik

using namespace fteik;

extern "C" int fteik3d_interp(const double *v, int n0, int n1, int n2, int stride,
                              const double *origin, const double *d, const double *q, int nq,
                              double fill, double *out, void *stream) {
    const int n[3] = {n0, n1, n2};
    if (!v || !q || !out || !origin || !d || stride < 1 || nq < 0 || !axes_ok(n, d, 3))
        return FTEIK_ERR_ARG;
    if (nq == 0) return 0;
    Axis X{origin[0], d[0], n0}, Y{origin[1], d[1], n1}, Z{origin[2], d[2], n2};
    hipLaunchKernelGGL(interp3d_kernel, dim3((nq + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       v, X, Y, Z, stride, q, nq, fill, out);
    return hip_status(hipGetLastError());
}

extern "C" int fteik3d_vinterp(const double *tt, int n0, int n1, int n2, const double *origin,
                               const double *d, const double *q, int nq, const double *src,
                               double vzero, double fill, double *out, void *stream) {
    const int n[3] = {n0, n1, n2};
    if (!tt || !q || !out || !origin || !d || !src || nq < 0 || !axes_ok(n, d, 3))
        return FTEIK_ERR_ARG;
    if (nq == 0) return 0;
    Axis X{origin[0], d[0], n0}, Y{origin[1], d[1], n1}, Z{origin[2], d[2], n2};
    hipLaunchKernelGGL(vinterp3d_kernel, dim3((nq + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       tt, X, Y, Z, q, nq, src[0], src[1], src[2], vzero, fill, out);
    return hip_status(hipGetLastError());
}

extern "C" int fteik2d_interp(const double *v, int n0, int n1, int stride, const double *origin,
                              const double *d, const double *q, int nq, double fill, double *out,
                              void *stream) {
    const int n[2] = {n0, n1};
    if (!v || !q || !out || !origin || !d || stride < 1 || nq < 0 || !axes_ok(n, d, 2))
        return FTEIK_ERR_ARG;
    if (nq == 0) return 0;
    Axis X{origin[0], d[0], n0}, Y{origin[1], d[1], n1};
    hipLaunchKernelGGL(interp2d_kernel, dim3((nq + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       v, X, Y, stride, q, nq, fill, out);
    return hip_status(hipGetLastError());
}

extern "C" int fteik2d_vinterp(const double *tt, int n0, int n1, const double *origin,
                               const double *d, const double *q, int nq, const double *src,
                               double vzero, double fill, double *out, void *stream) {
    const int n[2] = {n0, n1};
    if (!tt || !q || !out || !origin || !d || !src || nq < 0 || !axes_ok(n, d, 2))
        return FTEIK_ERR_ARG;
    if (nq == 0) return 0;
    Axis X{origin[0], d[0], n0}, Y{origin[1], d[1], n1};
    hipLaunchKernelGGL(vinterp2d_kernel, dim3((nq + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       tt, X, Y, q, nq, src[0], src[1], vzero, fill, out);
    return hip_status(hipGetLastError());
}
