// Device-side grid axes and (apparent-velocity) bi/trilinear interpolation shared by the
// interpolation and ray-tracing kernels.
#pragma once

#include "fteik_device.hpp"

namespace fteik {

// A regular axis: value(i) = origin + d * i, evaluated exactly like the reference's
// `origin + gridsize * np.arange(n)` (fteikpy/_base.py:125-133, :214-227).
struct Axis {
    double o, d;
    int n;
    __device__ __forceinline__ double at(int i) const { return o + d * (double)i; }
    __device__ __forceinline__ double last() const { return at(n - 1); }
    __device__ __forceinline__ bool contains(double q) const { return at(0) <= q && q <= last(); }
    // np.searchsorted(axis, q, side="right") - 1 for at(0) <= q: a floor() guess corrected
    // against the exact axis values, so the result matches the reference's binary search.
    __device__ __forceinline__ int cell(double q) const {
        double g = floor((q - o) / d);
        int i = g < 0.0 ? 0 : (g > (double)(n - 1) ? n - 1 : (int)g);
        while (i + 1 < n && at(i + 1) <= q) ++i;
        while (i > 0 && at(i) > q) --i;
        return i;
    }
};

__device__ __forceinline__ double dist3(double x1, double y1, double z1, double x2, double y2,
                                        double z2) {
    const double a = x1 - x2, b = y1 - y2, c = z1 - z2;
    return sqrt(a * a + b * b + c * c);
}
__device__ __forceinline__ double dist2(double x1, double y1, double x2, double y2) {
    const double a = x1 - x2, b = y1 - y2;
    return sqrt(a * a + b * b);
}

// Cell corners around a query; on the last node the missing neighbour is a ghost node at
// 2*x[-1] - x[-2] (fteikpy/_interp/_interp3d.py:28-162).
struct Corner {
    int i1;
    bool has2;
    double x1, x2;
};
__device__ __forceinline__ Corner corner(const Axis &a, int i1) {
    Corner c;
    c.i1 = i1;
    c.has2 = i1 != a.n - 1;
    c.x1 = a.at(i1);
    c.x2 = c.has2 ? a.at(i1 + 1) : 2.0 * c.x1 - a.at(a.n - 2);
    return c;
}

// Trilinear interpolation of v (strided element access), fteikpy/_interp/_interp3d.py:7-174.
__device__ inline double interp3(const Axis &X, const Axis &Y, const Axis &Z, const double *v,
                                 int stride, double xq, double yq, double zq, double fill) {
    if (!(X.contains(xq) && Y.contains(yq) && Z.contains(zq))) return fill;
    const Corner cx = corner(X, X.cell(xq)), cy = corner(Y, Y.cell(yq)), cz = corner(Z, Z.cell(zq));
    auto V = [&](int a, int b, int c) {
        return v[(((long long)a * Y.n + b) * Z.n + c) * stride];
    };
    const int i = cx.i1, j = cy.i1, k = cz.i1;
    const double v111 = V(i, j, k);
    const double v211 = cx.has2 ? V(i + 1, j, k) : 1.0;
    const double v121 = cy.has2 ? V(i, j + 1, k) : 1.0;
    const double v221 = (cx.has2 && cy.has2) ? V(i + 1, j + 1, k) : 1.0;
    const double v112 = cz.has2 ? V(i, j, k + 1) : 1.0;
    const double v212 = (cx.has2 && cz.has2) ? V(i + 1, j, k + 1) : 1.0;
    const double v122 = (cy.has2 && cz.has2) ? V(i, j + 1, k + 1) : 1.0;
    const double v222 = (cx.has2 && cy.has2 && cz.has2) ? V(i + 1, j + 1, k + 1) : 1.0;
    const double x1 = cx.x1, x2 = cx.x2, y1 = cy.x1, y2 = cy.x2, z1 = cz.x1, z2 = cz.x2;
    double vq = v111 * fabs((x2 - xq) * (y2 - yq) * (z2 - zq));
    vq += v211 * fabs((x1 - xq) * (y2 - yq) * (z2 - zq));
    vq += v121 * fabs((x2 - xq) * (y1 - yq) * (z2 - zq));
    vq += v221 * fabs((x1 - xq) * (y1 - yq) * (z2 - zq));
    vq += v112 * fabs((x2 - xq) * (y2 - yq) * (z1 - zq));
    vq += v212 * fabs((x1 - xq) * (y2 - yq) * (z1 - zq));
    vq += v122 * fabs((x2 - xq) * (y1 - yq) * (z1 - zq));
    vq += v222 * fabs((x1 - xq) * (y1 - yq) * (z1 - zq));
    vq /= fabs((x2 - x1) * (y2 - y1) * (z2 - z1));
    return vq;
}

// Bilinear interpolation, fteikpy/_interp/_interp2d.py:7-74.
__device__ inline double interp2(const Axis &X, const Axis &Y, const double *v, int stride,
                                 double xq, double yq, double fill) {
    if (!(X.contains(xq) && Y.contains(yq))) return fill;
    const Corner cx = corner(X, X.cell(xq)), cy = corner(Y, Y.cell(yq));
    auto V = [&](int a, int b) { return v[((long long)a * Y.n + b) * stride]; };
    const int i = cx.i1, j = cy.i1;
    const double v11 = V(i, j);
    const double v21 = cx.has2 ? V(i + 1, j) : 1.0;
    const double v12 = cy.has2 ? V(i, j + 1) : 1.0;
    const double v22 = (cx.has2 && cy.has2) ? V(i + 1, j + 1) : 1.0;
    const double x1 = cx.x1, x2 = cx.x2, y1 = cy.x1, y2 = cy.x2;
    double vq = v11 * fabs((x2 - xq) * (y2 - yq));
    vq += v21 * fabs((x1 - xq) * (y2 - yq));
    vq += v12 * fabs((x2 - xq) * (y1 - yq));
    vq += v22 * fabs((x1 - xq) * (y1 - yq));
    vq /= fabs((x2 - x1) * (y2 - y1));
    return vq;
}

}  // namespace fteik
